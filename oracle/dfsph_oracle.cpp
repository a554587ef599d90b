// DFSPH step oracle — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference's simulation step, used by the tests,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg to check the HIP step.
// The product path (fluid-simulation-app_amd/csrc) never links, loads or
// calls anything here.
//
// What it restates (paths relative to the reference repository):
//   grid        crates/fluid-core/src/sim_grid.wgsl:36-72 (cell_of, count,
//               scatter), sim_scan.wgsl:97-120 (scan_single), sim.rs:38-43
//   solver      crates/fluid-core/src/sim_solve.wgsl:103-756
//   stats       sim_solve.wgsl:800-874
//   tracers     crates/fluid-core/src/sim_tracers.wgsl:54-183
//   order       crates/fluid-core/src/render.rs:3175-3203 (dispatch order),
//               render.rs:77-79 (refine passes)
//
// Ordering choices (the reference leaves them open):
//   * The reference's scatter takes any in-cell order (an atomic cursor,
//     sim_grid.wgsl:58). The oracle takes the stable one: within a cell,
//     particles keep the order of the resting set they came from.
//   * The reference's sweeps split each particle over 8 lanes and fold
//     (sim_solve.wgsl:377, 398-402) and append the neighbour list through a
//     workgroup atomic (sim_solve.wgsl:463). The oracle's density sweep is the
//     one serial walk of the 27 cells in the reference's cell numbering (x
//     fastest, then y, then z, sim_solve.wgsl:436-438), particles in slot
//     order within a cell, and the list in that same order. Every sweep's
//     pair sums (density_div, div_apply, forces_eval, forces_den_apply,
//     den_kappa, den_apply) split the neighbours in that order over 4 lanes
//     with the reference's strided assignment and xor-butterfly fold
//     (kListLanes, lane_fold below).
//   * forces_eval divides the pair vector x by max(r, 1e-5) and by
//     max(r, 1e-9) (sim_solve.wgsl:559, 567-568); the oracle multiplies x by
//     one reciprocal of r instead (the two maxima agree for r >= 1e-5), so
//     each quotient is the product with the correctly rounded reciprocal.
//   * Per-substep constants the reference evaluates per invocation with
//     exp() (XSPH blend, lens chase, tracer charge decay) are evaluated once
//     per substep with the host expf/exp.
//
// A line-faithful mode (oracle_set_faithful) takes the reference's own
// arithmetic for the three choices above that differ from it: the density
// walk of LANES = 8 lanes over strided stencil cells with their lane sums
// and xor-butterfly fold (sim_solve.wgsl:377, 398-402, 436-471), the list
// appended in the lanes' lockstep order and split over 8 lanes in every
// later sweep, x / max(r, eps) as true quotients (sim_solve.wgsl:559,
// 567-568), exp() in the fingers (sim_solve.wgsl:124), and the 256-thread
// strided sum of reduce_stats (sim_solve.wgsl:807-815). The HIP step is
// held to it within the per-substep tolerance, and to the default mode bit
// for bit.
//
// Every floating-point expression keeps the reference's operand order, and
// the file is compiled with -ffp-contract=off, so float results are those
// of IEEE single precision evaluated in that order.
//
// The double instantiation (the FP64 oracle) runs the same method in double
// precision. Its grid is built from the positions rounded to float, so that
// both instantiations sort the same input identically.

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

// A forward-mode dual number in double precision: a value and its
// derivative along one seeded direction. Comparisons and branches follow the
// value, so a dual run takes exactly the double run's path and its tangent is
// the exact derivative of that code (the FP64 reference gradient along the
// seed: initial velocities, viscosity or a substep's gravity).
struct Dual {
    double v = 0.0, d = 0.0;
    Dual() = default;
    Dual(double x) : v(x), d(0.0) {}  // NOLINT: constants and parameters
    Dual(double x, double dx) : v(x), d(dx) {}
    explicit operator float() const { return (float)v; }
    explicit operator double() const { return v; }
    Dual& operator+=(const Dual& b) { d += b.d; v += b.v; return *this; }
    Dual& operator-=(const Dual& b) { d -= b.d; v -= b.v; return *this; }
    Dual& operator*=(const Dual& b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
    Dual& operator/=(const Dual& b) {
        const double q = v / b.v;
        d = (d - q * b.d) / b.v;
        v = q;
        return *this;
    }
};
inline Dual operator+(Dual a, const Dual& b) { return a += b; }
inline Dual operator-(Dual a, const Dual& b) { return a -= b; }
inline Dual operator*(Dual a, const Dual& b) { return a *= b; }
inline Dual operator/(Dual a, const Dual& b) { return a /= b; }
inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
inline bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }
inline Dual dsqrt(const Dual& a) {
    const double r = std::sqrt(a.v);
    return Dual(r, r > 0.0 ? a.d / (2.0 * r) : 0.0);
}
inline Dual dexp(const Dual& a) {
    const double e = std::exp(a.v);
    return Dual(e, e * a.d);
}

// A forward-mode dual number whose value is single precision and whose
// derivative is double: every value is computed by exactly the float
// operation the float instantiation performs (so a run takes the FP32 step's
// path and produces its bits), and the derivative is that of those float
// operations, carried in double. Its tangent is the exact derivative of the
// FP32 step itself (the trajectory the GPU runs), which the GPU adjoint, a
// float reverse sweep over the same trajectory, must reproduce.
struct DualF {
    float v = 0.0f;
    double d = 0.0;
    DualF() = default;
    DualF(double x) : v((float)x), d(0.0) {}  // NOLINT: constants and parameters, rounded as float literals are
    DualF(float x, double dx) : v(x), d(dx) {}
    explicit operator float() const { return v; }
    explicit operator double() const { return (double)v; }
    DualF& operator+=(const DualF& b) { d += b.d; v += b.v; return *this; }
    DualF& operator-=(const DualF& b) { d -= b.d; v -= b.v; return *this; }
    DualF& operator*=(const DualF& b) { d = d * (double)b.v + (double)v * b.d; v *= b.v; return *this; }
    DualF& operator/=(const DualF& b) {
        const float q = v / b.v;
        d = (d - (double)q * b.d) / (double)b.v;
        v = q;
        return *this;
    }
};
inline DualF operator+(DualF a, const DualF& b) { return a += b; }
inline DualF operator-(DualF a, const DualF& b) { return a -= b; }
inline DualF operator*(DualF a, const DualF& b) { return a *= b; }
inline DualF operator/(DualF a, const DualF& b) { return a /= b; }
inline DualF operator-(const DualF& a) { return DualF(-a.v, -a.d); }
inline bool operator<(const DualF& a, const DualF& b) { return a.v < b.v; }
inline bool operator>(const DualF& a, const DualF& b) { return a.v > b.v; }
inline bool operator<=(const DualF& a, const DualF& b) { return a.v <= b.v; }
inline bool operator>=(const DualF& a, const DualF& b) { return a.v >= b.v; }
inline bool operator==(const DualF& a, const DualF& b) { return a.v == b.v; }
inline bool operator!=(const DualF& a, const DualF& b) { return a.v != b.v; }

namespace {

// Whether T computes in single precision (float, or DualF's value).
template <typename T> struct IsF32 : std::is_same<T, float> {};
template <> struct IsF32<DualF> : std::true_type {};

// Sets the tangent of a dual value (nothing at float or double).
inline void seed(float&, double) {}
inline void seed(double&, double) {}
inline void seed(Dual& x, double d) { x.d = d; }
inline void seed(DualF& x, double d) { x.d = d; }
inline double tangent(float) { return 0.0; }
inline double tangent(double) { return 0.0; }
inline double tangent(const Dual& x) { return x.d; }
inline double tangent(const DualF& x) { return x.d; }
inline float valf(float x) { return x; }
inline float valf(double x) { return (float)x; }
inline float valf(const Dual& x) { return (float)x.v; }
inline float valf(const DualF& x) { return x.v; }

// Tolerance mode's per-particle error e as a 2^-32 fixed-point integer,
// saturated at 2^40 (e = 256), NaN and non-positive errors zero: the HIP
// step's fixed_err (solve.hip), whose integer sums are order-free.
inline uint64_t fixed_err(float e) {
    e = e * 4294967296.0f;
    return e > 0.0f ? (uint64_t)(e < 1099511627776.0f ? e : 1099511627776.0f) : 0u;
}

// sqrt and exp at precision T (the dual forms carry the derivative).
inline float tsqrt(float x) { return std::sqrt(x); }
inline double tsqrt(double x) { return std::sqrt(x); }
inline Dual tsqrt(const Dual& x) { return dsqrt(x); }
inline DualF tsqrt(const DualF& x) {
    const float r = std::sqrt(x.v);
    return DualF(r, r > 0.0f ? x.d / (2.0 * (double)r) : 0.0);
}
inline float texp(float x) { return std::exp(x); }
inline double texp(double x) { return std::exp(x); }
inline Dual texp(const Dual& x) { return dexp(x); }
inline DualF texp(const DualF& x) {
    const float e = std::exp(x.v);
    return DualF(e, (double)e * x.d);
}

struct Params {  // SimParams, sim.rs:325-350 / sim_solve.wgsl:11-23
    float box_min[3];
    float cell;
    uint32_t dims[3];
    uint32_t count;
    float h, mass, rho0, inv_h, sigma, sigma_h, support_sq, pad;
};
static_assert(sizeof(Params) == 64, "SimParams is 64 bytes");

struct StepBlock {  // Step, sim.rs:200-252 / sim_solve.wgsl:29-41
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

// Physics constants, sim_solve.wgsl:84-159.
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
// Tracer constants, sim_tracers.wgsl:39-52.
constexpr float FIXED = 65536.0f;
constexpr float TAU = 3.0f;
constexpr float TAU_STRAY = 0.25f;
constexpr float T_CHARGE = 4.0f;
constexpr float HUGE_F = 1e30f;

// ---- integer helpers --------------------------------------------------

// WGSL/Rust float -> u32: truncate, negatives and NaN to 0, saturate.
inline uint32_t sat_u32(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}
inline int32_t sat_i32(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)f;
}

// sim_grid.wgsl:36-42, sim.rs:38-43 — always in float.
inline uint32_t cell_of(const Params& p, float x, float y, float z) {
    uint32_t cx = std::min(sat_u32((x - p.box_min[0]) / p.cell), p.dims[0] - 1u);
    uint32_t cy = std::min(sat_u32((y - p.box_min[1]) / p.cell), p.dims[1] - 1u);
    uint32_t cz = std::min(sat_u32((z - p.box_min[2]) / p.cell), p.dims[2] - 1u);
    return (cz * p.dims[1] + cy) * p.dims[0] + cx;
}

// sim_tracers.wgsl:54-58
inline uint32_t pcg(uint32_t x) {
    uint32_t h = x * 747796405u + 2891336453u;
    h = ((h >> ((h >> 28u) + 4u)) ^ h) * 277803737u;
    return (h >> 22u) ^ h;
}

// IEEE binary16 bits of a float, round to nearest even (pack2x16float).
inline uint16_t f32_to_f16_rne(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) {  // inf / nan
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    }
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds past 65504
    if (ax < 0x38800000u) {  // below half's smallest normal: subnormal grid
        if (ax < 0x33000000u) return (uint16_t)sign;  // < 2^-25 rounds to 0
        uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
        uint32_t e = ax >> 23;                         // 102..112
        uint32_t shift = 126u - e;                     // 14..24
        uint32_t q = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (q & 1u))) q++;
        return (uint16_t)(sign | q);
    }
    uint32_t q = (ax - 0x38000000u) >> 13;  // rebias 127 -> 15
    uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
    return (uint16_t)(sign | q);
}
inline float f16_to_f32(uint16_t hbits) {
    uint32_t sign = (uint32_t)(hbits & 0x8000u) << 16;
    uint32_t e = (hbits >> 10) & 0x1fu;
    uint32_t m = hbits & 0x3ffu;
    uint32_t out;
    if (e == 0) {
        if (m == 0) {
            out = sign;
        } else {
            float v = (float)m * (1.0f / 16777216.0f);  // m * 2^-24, exact
            std::memcpy(&out, &v, 4);
            out |= sign;
        }
    } else if (e == 31) {
        out = sign | 0x7f800000u | (m << 13);
    } else {
        out = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &out, 4);
    return f;
}
// pack2x16unorm: floor(0.5 + 65535 * clamp(v, 0, 1))
inline uint32_t unorm16(float v) {
    float c = std::min(std::max(v, 0.0f), 1.0f);
    return (uint32_t)(0.5f + 65535.0f * c);
}

// ---- templated floating-point restatement -------------------------------

// A literal at precision T: float instantiations take the f-suffixed literal
// (rounded once, as a WGSL f32 literal is), double ones the double literal.
#define K(x) (IsF32<T>::value ? (T)(x##f) : (T)(x))

template <typename T>
struct V3 {
    T x, y, z;
};
template <typename T> inline V3<T> v3(T x, T y, T z) { return {x, y, z}; }
template <typename T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> inline T tmax(T a, T b) { return a > b ? a : b; }  // WGSL max, no NaN
template <typename T> inline T tmin(T a, T b) { return a < b ? a : b; }
template <typename T> inline T mix(T a, T b, T t) { return a * (T(1) - t) + b * t; }

// The reference's compile-time physical coefficients as runtime values
// (sim_solve.wgsl:135-137, 161-172, 657): viscosity, and switches for the
// terms the adjoint's differentiable profile leaves out. Defaults are the
// reference's.
struct Physics {
    float mu = DYNAMIC_VISCOSITY;
    float k_near = K_NEAR;
    float tension_scale = 1.0f;
    float adhesion_scale = 1.0f;
    float xsph_rate = XSPH_RATE;
};

// Constants of the method at precision T: parameter values come from the
// f32 SimParams block; const-expressions of the shader are folded at T.
template <typename T>
struct Consts {
    T box_min[3], cell, h, mass, rho0, inv_h, sigma, sigma_h, support_sq;
    T wall_lo[3], wall_hi[3];
    T coh_c, coh_k, coh_near, gamma, beta;
    T visc2, heat_cond, heat_visc, eps_h2, k_near;
    explicit Consts(const Params& p, const Physics& phys = Physics()) {
        for (int i = 0; i < 3; ++i) box_min[i] = (T)p.box_min[i];
        cell = (T)p.cell;
        h = (T)p.h;
        mass = (T)p.mass;
        rho0 = (T)p.rho0;
        inv_h = (T)p.inv_h;
        sigma = (T)p.sigma;
        sigma_h = (T)p.sigma_h;
        support_sq = (T)p.support_sq;
        for (int i = 0; i < 3; ++i) {  // wall_lo = box_min + cell, sim_solve.wgsl:237-239
            wall_lo[i] = box_min[i] + cell;
            wall_hi[i] = -wall_lo[i];
        }
        // forces_eval hoists, sim_solve.wgsl:534-537
        coh_c = T(2) * h;
        T c3 = coh_c * coh_c * coh_c;
        coh_k = T(32) / (K(3.14159265) * c3 * c3 * c3);
        coh_near = c3 * c3 / T(64);
        // tension_gamma / adhesion_beta, sim_solve.wgsl:161-172
        T c = T(2) * h;
        gamma = (T(SIGMA_WATER) * T(7040)) / (T(21) * T(LATTICE_SIGMA) * rho0 * rho0 * c * c) *
                T(phys.tension_scale);
        beta = (T(SIGMA_WATER) * (T(1) + T(COS_CONTACT)) * K(1558.4)) /
               (T(LATTICE_SIGMA) * rho0 * rho0 * c * c) * T(phys.adhesion_scale);
        // sim_solve.wgsl:559-565 constant factors
        visc2 = T(2) * T(phys.mu);
        heat_cond = T(2) * T(CONDUCTIVITY) / T(HEAT_CAPACITY);
        heat_visc = K(0.5) * T(phys.mu) / T(HEAT_CAPACITY);
        k_near = T(phys.k_near);
        eps_h2 = K(0.01) * h * h;
    }
};

// The list sweeps split each particle's neighbour list over kListLanes
// lanes, as the reference splits it over LANES = 8 (sim_solve.wgsl:377,
// 546, 601, 684): lane l takes entries l, l + kListLanes, ... in order, and
// the lane partials fold the way lane_fold does (sim_solve.wgsl:398-402),
// xor 2 then xor 1, so lane 0 holds (p0 + p2) + (p1 + p3).
//
// The line-faithful mode (Sim::faithful) splits over the reference's own
// LANES = 8 and folds xor 4, xor 2, xor 1: lane 0 holds
// ((p0 + p4) + (p2 + p6)) + ((p1 + p5) + (p3 + p7)).
constexpr uint32_t kListLanes = 4;
constexpr uint32_t kRefLanes = 8;
template <typename A> inline A lane_fold(const A* p, uint32_t lanes) {
    if (lanes == kRefLanes) {
        const A a0 = p[0] + p[4], a1 = p[1] + p[5], a2 = p[2] + p[6], a3 = p[3] + p[7];
        return (a0 + a2) + (a1 + a3);
    }
    const A t0 = p[0] + p[2];
    const A t1 = p[1] + p[3];
    return t0 + t1;
}

// sim_solve.wgsl:176-182
template <typename T> inline T kernel_q(const Consts<T>& k, T q) {
    if (q < T(1)) return k.sigma * (T(1) - K(1.5) * q * q * (T(1) - K(0.5) * q));
    T t = T(2) - q;
    return k.sigma * K(0.25) * t * t * t;
}
// sim_solve.wgsl:184-196
template <typename T> inline V3<T> grad_q(const Consts<T>& k, V3<T> x, T r, T q) {
    if (r < K(1e-9)) return v3(T(0), T(0), T(0));
    T dw;
    if (q < T(1)) {
        dw = k.sigma_h * (T(-3) * q + K(2.25) * q * q);
    } else {
        T t = T(2) - q;
        dw = k.sigma_h * (K(-0.75) * t * t);
    }
    return (dw / r) * x;
}
// sim_solve.wgsl:201-210
template <typename T> inline T cohesion(T r, T c, T kk, T nearc) {
    if (r >= c || r < K(1e-9)) return T(0);
    T a = (c - r) * (c - r) * (c - r) * r * r * r;
    if (T(2) * r > c) return kk * a;
    return kk * (T(2) * a - nearc);
}
// sim_solve.wgsl:213-222 (= sim.rs:261-272)
template <typename T> inline T wall_density(T t) {
    if (t >= T(2)) return T(0);
    T t3 = t * t * t;
    if (t < T(1)) return K(0.5) - K(0.7) * t + t3 * (T(1.0 / 3.0) + t * t * (K(-0.15) + K(0.05) * t));
    return T(8.0 / 15.0) - K(0.8) * t + t3 * (T(2.0 / 3.0) + t * (K(-0.5) + t * (K(0.15) - t / T(60))));
}
// sim_solve.wgsl:226-235 (= sim.rs:279-289)
template <typename T> inline T wall_gradient(T t) {
    if (t >= T(2)) return T(0);
    T t2 = t * t;
    if (t < T(1)) return K(0.7) + t2 * (T(-1) + t2 * (K(0.75) - t * K(0.3)));
    return K(0.8) + t2 * (T(-2) + t * (T(2) + t * (K(-0.75) + t * K(0.1))));
}
// sim_solve.wgsl:249-270 — the fitted coefficient table is the reference's.
template <typename T> inline T wedge(T t1, T t2) {
    if (t1 * t1 + t2 * t2 >= T(4)) return T(0);
    const T a0 = K(2.5005807e-01) + t2 * (K(-3.4681153e-01) + t2 * (K(-4.0986882e-02) + t2 * (K(3.0301620e-01) + t2 * (K(-1.9202736e-01) + t2 * (K(4.6947582e-02) + t2 * K(-3.8543515e-03))))));
    const T a1 = K(-3.4681153e-01) + t2 * (K(4.8587248e-01) + t2 * (K(1.0092180e-02) + t2 * (K(-2.9978363e-01) + t2 * (K(1.5320335e-01) + t2 * K(-2.2544282e-02)))));
    const T a2 = K(-4.0986882e-02) + t2 * (K(1.0092180e-02) + t2 * (K(-5.4145610e-02) + t2 * (K(7.9597369e-02) + t2 * K(-2.3350373e-02))));
    const T a3 = K(3.0301620e-01) + t2 * (K(-2.9978363e-01) + t2 * (K(7.9597369e-02) + t2 * K(-8.2431946e-03)));
    const T a4 = K(-1.9202736e-01) + t2 * (K(1.5320335e-01) + t2 * K(-2.3350373e-02));
    const T a5 = K(4.6947582e-02) + t2 * K(-2.2544282e-02);
    const T a6 = K(-3.8543515e-03);
    return a0 + t1 * (a1 + t1 * (a2 + t1 * (a3 + t1 * (a4 + t1 * (a5 + t1 * a6)))));
}
// sim_solve.wgsl:272-293
template <typename T> inline T wedge_d(T t1, T t2) {
    if (t1 * t1 + t2 * t2 >= T(4)) return T(0);
    const T b0 = K(-3.4611994e-01) + t2 * (K(4.4757673e-01) + t2 * (K(1.1932939e-01) + t2 * (K(-4.6460101e-01) + t2 * (K(2.7637253e-01) + t2 * (K(-6.2846025e-02) + t2 * K(4.3871652e-03))))));
    const T b1 = K(-6.3622823e-02) + t2 * (K(2.0521006e-01) + t2 * (K(-3.7311339e-01) + t2 * (K(3.7508155e-01) + t2 * (K(-1.6629713e-01) + t2 * K(2.5030129e-02)))));
    const T b2 = K(7.8352497e-01) + t2 * (K(-1.2658344e+00) + t2 * (K(5.8898945e-01) + t2 * (K(-1.0772903e-01) + t2 * K(1.3535777e-02))));
    const T b3 = K(-4.9470084e-01) + t2 * (K(9.1302071e-01) + t2 * (K(-3.4330355e-01) + t2 * K(2.4034609e-02)));
    const T b4 = K(-2.5605481e-02) + t2 * (K(-2.0067981e-01) + t2 * K(5.9854468e-02));
    const T b5 = K(8.8898623e-02) + t2 * K(3.9914838e-03);
    const T b6 = K(-1.7788321e-02);
    return b0 + t1 * (b1 + t1 * (b2 + t1 * (b3 + t1 * (b4 + t1 * (b5 + t1 * b6)))));
}
// sim_solve.wgsl:334-340
template <typename T> inline T wall_adhesion(T t) {
    T u = tmin(tmax(T(2) - t, T(0)), T(2));
    return u * u * (K(1.7847015e-3) + u * (K(3.3392196e-3) + u * (K(-4.3394677e-3) + u * (K(1.6711868e-3) + u * K(-2.1788750e-4)))));
}
// sim_solve.wgsl:295-304
template <typename T> inline T wedge_dsum(T t, T a, T b, T c, T d) {
    return wedge_d(t, a) + wedge_d(t, b) + wedge_d(t, c) + wedge_d(t, d);
}
template <typename T> inline T wedge_sum(V3<T> tl, V3<T> th) {
    T w = wedge(tl.x, tl.y) + wedge(tl.x, th.y) + wedge(th.x, tl.y) + wedge(th.x, th.y);
    w += wedge(tl.x, tl.z) + wedge(tl.x, th.z) + wedge(th.x, tl.z) + wedge(th.x, th.z);
    w += wedge(tl.y, tl.z) + wedge(tl.y, th.z) + wedge(th.y, tl.z) + wedge(th.y, th.z);
    return w;
}
// sim_solve.wgsl:310-328 (note: 1/h evaluated here, as the shader does)
template <typename T> inline V3<T> wall_grad_sum(const Consts<T>& k, V3<T> pos) {
    V3<T> lo = v3(k.wall_lo[0], k.wall_lo[1], k.wall_lo[2]);
    V3<T> hi = v3(k.wall_hi[0], k.wall_hi[1], k.wall_hi[2]);
    T inv_h = T(1) / k.h;
    V3<T> tl = inv_h * (pos - lo);
    V3<T> th = inv_h * (hi - pos);
    T scale = k.rho0 * inv_h;
    V3<T> g = v3(T(0), T(0), T(0));
    g.x -= scale * (wall_gradient(tl.x) + wedge_dsum(tl.x, tl.y, th.y, tl.z, th.z));
    g.x += scale * (wall_gradient(th.x) + wedge_dsum(th.x, tl.y, th.y, tl.z, th.z));
    g.y -= scale * (wall_gradient(tl.y) + wedge_dsum(tl.y, tl.x, th.x, tl.z, th.z));
    g.y += scale * (wall_gradient(th.y) + wedge_dsum(th.y, tl.x, th.x, tl.z, th.z));
    g.z -= scale * (wall_gradient(tl.z) + wedge_dsum(tl.z, tl.x, th.x, tl.y, th.y));
    g.z += scale * (wall_gradient(th.z) + wedge_dsum(th.z, tl.x, th.x, tl.y, th.y));
    return g;
}
// sim_solve.wgsl:349-361
template <typename T> inline V3<T> wall_adh_sum(const Consts<T>& k, V3<T> pos) {
    T inv_h = T(1) / k.h;
    V3<T> a = v3(T(0), T(0), T(0));
    a.x -= wall_adhesion((pos.x - k.wall_lo[0]) * inv_h);
    a.x += wall_adhesion((k.wall_hi[0] - pos.x) * inv_h);
    a.y -= wall_adhesion((pos.y - k.wall_lo[1]) * inv_h);
    a.y += wall_adhesion((k.wall_hi[1] - pos.y) * inv_h);
    a.z -= wall_adhesion((pos.z - k.wall_lo[2]) * inv_h);
    a.z += wall_adhesion((k.wall_hi[2] - pos.z) * inv_h);
    return a;
}

// Per-substep uniforms (Step block plus the host-evaluated exponentials).
template <typename T>
struct StepT {
    V3<T> force, omega, domega;
    T dt, v_clamp, touch_r;
    uint32_t touch_n, seed;
    T touches[5][4];
    T c_xsph, chase;
    explicit StepT(const StepBlock& s, const Physics& phys = Physics()) {
        force = v3((T)s.force[0], (T)s.force[1], (T)s.force[2]);
        omega = v3((T)s.omega[0], (T)s.omega[1], (T)s.omega[2]);
        domega = v3((T)s.domega[0], (T)s.domega[1], (T)s.domega[2]);
        dt = (T)s.dt;
        v_clamp = (T)s.v_clamp;
        touch_r = (T)s.touch_r;
        touch_n = std::min<uint32_t>(s.touch_n, 5u);
        seed = s.seed;
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 4; ++j) touches[i][j] = (T)s.touches[i][j];
        c_xsph = T(1) - texp(-T(phys.xsph_rate) * dt);  // sim_solve.wgsl:657
        chase = T(1) - texp(-T(LENS_RATE) * dt);   // sim_solve.wgsl:741
    }
};

// e^x as a fixed sequence of float operations (the same sequence the HIP
// kernels run, so both sides round identically): Cody-Waite reduction with
// ln2 = 0.693145751953125 + 1.42860677e-06, a degree-7 Taylor polynomial,
// exact scaling by 2^k. The double oracle uses std::exp.
inline float exp_fixed(float x) {
    x = std::min(std::max(x, -87.0f), 88.0f);
    const float k = std::nearbyint(x * 1.44269504f);
    const float r = (x - k * 0.693145751953125f) - k * 1.42860677e-06f;
    const float p = 1.0f + r * (1.0f + r * (0.5f + r * (0.166666672f + r * (0.0416666679f + r * (0.00833333377f +
                    r * (0.00138888892f + r * 0.000198412701f))))));
    return std::ldexp(p, (int)k);
}
inline double exp_fixed(double x) { return std::exp(x); }
inline Dual exp_fixed(const Dual& x) { return dexp(x); }
inline DualF exp_fixed(const DualF& x) {
    const float e = exp_fixed(x.v);
    return DualF(e, (double)e * x.d);
}

// sim_solve.wgsl:103-127
template <typename T> inline V3<T> fingers(const StepT<T>& s, V3<T> pos, V3<T> v, bool faithful = false) {
    T pullx = 0, pully = 0, weight = 0;
    for (uint32_t i = 0; i < s.touch_n; ++i) {
        const T* t = s.touches[i];
        T dx = pos.x - t[0], dy = pos.y - t[1];
        T q2 = (dx * dx + dy * dy) / (s.touch_r * s.touch_r);
        if (q2 < T(1)) {
            T w = (T(1) - q2) * (T(1) - q2);
            pullx += w * t[2];
            pully += w * t[3];
            weight += w;
        }
    }
    if (weight <= T(0)) return v;
    // The faithful mode calls exp as the shader does; the kernels' fixed
    // sequence (exp_fixed) is the default.
    const T arg = -T(TOUCH_RATE) * weight * s.dt;
    T blend = T(1) - (faithful ? texp(arg) : exp_fixed(arg));
    return v3(mix(v.x, pullx / weight, blend), mix(v.y, pully / weight, blend), v.z);
}

template <typename T>
struct Sim {
    Params p;
    Consts<T> k;
    uint32_t n, cells, cap;
    // resting set (slot order of the last integrate)
    std::vector<T> pos, vel, pvel, pp, temp;  // pos/vel/pvel: 4 per particle
    std::vector<uint32_t> ids;                // identity carried through the sort
    // working set (cell order of this substep)
    std::vector<T> wpos, wvel, wpvel, wpp, wtemp;
    std::vector<uint32_t> wids;
    // scratch
    std::vector<T> density, alpha, kd, pressure, rho_near, accel, xsph, wgrad;
    std::vector<uint32_t> nbr, nbr_n;
    // grid
    std::vector<uint32_t> cellid, hist, starts, perm;
    uint64_t clamp_count = 0, nbr_over = 0;
    float stats[14] = {0};
    // tracers
    std::vector<int32_t> vel_grid, vel_flat;

    Physics phys;
    // Line-faithful mode: the reference's own arithmetic where the default
    // follows the HIP kernels' (LANES = 8 and the strided density walk,
    // true division in forces_eval, exp in the fingers, the 256-thread
    // strided stats sum).
    bool faithful = false;
    uint32_t lanes() const { return faithful ? kRefLanes : kListLanes; }
    // Tolerance mode (no counterpart in the reference, which fixes one
    // divergence pass and `refine` density passes, sim_solve.wgsl:665-703):
    // the density solve iterates until the mean of max(rho* - rho0, 0) /
    // rho0 is <= tol (min_passes .. max_passes passes), the divergence solve
    // until the mean of max(Drho/Dt, 0) dt / rho0 is <= div_tol (at most
    // div_max passes); the pass meeting the tolerance applies and is the
    // last. The means come from order-free fixed-point sums (fixed_err), so
    // the pass counts are the HIP step's exactly.
    float tol = 0.0f, div_tol = 0.0f;
    uint32_t min_passes = 1, max_passes = 0, div_max = 1;
    bool err_on = false;                 // density_finish / den_kappa fill err
    std::vector<uint64_t> err;           // per particle, this pass
    uint32_t last_passes[2] = {0, 0};    // density, divergence
    std::vector<uint64_t> pass_hist = std::vector<uint64_t>(2 * 257, 0);
    double mean_err() const {
        uint64_t sum = 0;
        for (uint32_t i = 0; i < n; ++i) sum += err[i];
        return (double)sum * (1.0 / 4294967296.0) / (double)(n ? n : 1u);
    }
    void set_physics(const Physics& ph) {
        phys = ph;
        k = Consts<T>(p, ph);
    }
    // Tangent seeds of the dual run: the next substep's gravity, viscosity.
    double dforce[3] = {0.0, 0.0, 0.0};
    void seed_mu(double dmu) {
        seed(k.visc2, 2.0 * dmu);
        seed(k.heat_visc, 0.5 * dmu / (double)HEAT_CAPACITY);
    }

    Sim(const Params& prm, uint32_t nbr_cap) : p(prm), k(prm) {
        n = p.count;
        cells = p.dims[0] * p.dims[1] * p.dims[2];
        cap = nbr_cap;
        auto r4 = [&](std::vector<T>& v) { v.assign((size_t)n * 4, T(0)); };
        auto r1 = [&](std::vector<T>& v) { v.assign((size_t)n, T(0)); };
        r4(pos); r4(vel); r4(pvel); r1(pp); r1(temp);
        r4(wpos); r4(wvel); r4(wpvel); r1(wpp); r1(wtemp);
        r1(density); r1(alpha); r1(kd); r1(pressure); r1(rho_near);
        r4(accel); r4(xsph); r4(wgrad);
        ids.resize(n); wids.resize(n);
        for (uint32_t i = 0; i < n; ++i) ids[i] = i;
        // The box starts at 293.15 K, all else at zero (render.rs:1540-1555).
        std::fill(temp.begin(), temp.end(), T(293.15));
        nbr.assign((size_t)n * cap, 0u);
        nbr_n.assign(n, 0u);
        cellid.assign(n, 0u);
        hist.assign(cells, 0u);
        starts.assign((size_t)cells + 1, 0u);
        perm.assign(n, 0u);
        vel_grid.assign((size_t)cells * 4, 0);
        vel_flat.assign((size_t)cells * 4, 0);
    }

    V3<T> wp(uint32_t i) const { return v3(wpos[4 * i], wpos[4 * i + 1], wpos[4 * i + 2]); }
    V3<T> wv(uint32_t i) const { return v3(wvel[4 * i], wvel[4 * i + 1], wvel[4 * i + 2]); }

    // count + scan_single + scatter, sim_grid.wgsl:44-72, sim_scan.wgsl:97-120
    void grid_build() {
        std::fill(hist.begin(), hist.end(), 0u);
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t c = cell_of(p, (float)pos[4 * i], (float)pos[4 * i + 1], (float)pos[4 * i + 2]);
            cellid[i] = c;
            hist[c]++;
        }
        uint32_t run = 0;
        for (uint32_t c = 0; c < cells; ++c) {
            starts[c] = run;
            run += hist[c];
        }
        starts[cells] = run;
        std::vector<uint32_t> cursor(starts.begin(), starts.end() - 1);
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t slot = cursor[cellid[i]]++;
            perm[slot] = i;
        }
        for (uint32_t s = 0; s < n; ++s) {
            uint32_t i = perm[s];
            for (int c = 0; c < 4; ++c) {
                wpos[4 * s + c] = pos[4 * i + c];
                wvel[4 * s + c] = vel[4 * i + c];
                wpvel[4 * s + c] = pvel[4 * i + c];
            }
            wpp[s] = pp[i];
            wtemp[s] = temp[i];
            wids[s] = ids[i];
        }
    }

    // sim_solve.wgsl:363-368 in float (the cell arithmetic is float)
    void cell_base(uint32_t i, int32_t out[3]) const {
        for (int a = 0; a < 3; ++a)
            out[a] = (int32_t)std::min(sat_u32(((float)wpos[4 * i + a] - p.box_min[a]) / p.cell), p.dims[a] - 1u);
    }

    // sim_solve.wgsl:411-518: the walk in the reference's cell order builds
    // the list; the pair sums split the neighbours over kListLanes lanes
    // (neighbour m, past the cap included, to lane m % kListLanes; lane 0
    // carries the self term, sim_solve.wgsl:426) and fold with lane_fold
    // (sim_solve.wgsl:470-471). The faithful mode walks as the reference's
    // lanes do instead (density_walk_faithful).
    void density_div(const StepT<T>& s) {
        if (faithful) {
            density_walk_faithful(s);
            return;
        }
        uint64_t over = 0;
#pragma omp parallel for schedule(static) reduction(+ : over)
        for (int64_t pi = 0; pi < (int64_t)n; ++pi) {
            uint32_t pidx = (uint32_t)pi;
            V3<T> pos_i = wp(pidx), vel_i = wv(pidx);
            T w_l[kListLanes] = {}, rn_l[kListLanes] = {}, gsq_l[kListLanes] = {}, drho_l[kListLanes] = {};
            V3<T> gsum_l[kListLanes];
            for (uint32_t l = 0; l < kListLanes; ++l) gsum_l[l] = v3(T(0), T(0), T(0));
            w_l[0] = k.sigma;
            uint32_t found = 0;
            int32_t base[3];
            cell_base(pidx, base);
            for (int cell = 0; cell < 27; ++cell) {
                int32_t cx = base[0] + (cell % 3) - 1;
                int32_t cy = base[1] + (cell / 3) % 3 - 1;
                int32_t cz = base[2] + cell / 9 - 1;
                if (cx < 0 || cy < 0 || cz < 0 || cx >= (int32_t)p.dims[0] || cy >= (int32_t)p.dims[1] ||
                    cz >= (int32_t)p.dims[2])
                    continue;
                uint32_t c = ((uint32_t)cz * p.dims[1] + (uint32_t)cy) * p.dims[0] + (uint32_t)cx;
                uint32_t end = starts[c + 1];
                for (uint32_t kk = starts[c]; kk < end; ++kk) {
                    if (kk == pidx) continue;
                    V3<T> x = pos_i - wp(kk);
                    T r2 = dot(x, x);
                    if (r2 >= k.support_sq) continue;
                    const uint32_t l = found % kListLanes;
                    density_pair(vel_i, kk, x, r2, w_l[l], rn_l[l], gsq_l[l], drho_l[l], gsum_l[l]);
                    if (found < cap) nbr[(size_t)pidx * cap + found] = kk;
                    found++;
                }
            }
            nbr_n[pidx] = std::min(found, cap);
            if (found > cap) over += found - cap;
            density_finish(s, pidx, pos_i, vel_i, lane_fold(w_l, kListLanes), lane_fold(rn_l, kListLanes),
                           lane_fold(gsq_l, kListLanes), lane_fold(drho_l, kListLanes),
                           lane_fold(gsum_l, kListLanes));
        }
        nbr_over += over;
    }

    // One accepted pair's terms of density_div (sim_solve.wgsl:448-460).
    inline void density_pair(V3<T> vel_i, uint32_t kk, V3<T> x, T r2, T& w, T& rn, T& gsq, T& drho,
                             V3<T>& gsum) const {
        T r = tsqrt(r2);
        T q = r * k.inv_h;
        w += kernel_q(k, q);
        T wn = tmax(T(1) - q, T(0));
        rn += wn * wn * wn;
        V3<T> gw = grad_q(k, x, r, q);
        gsum = gsum + gw;
        gsq += dot(gw, gw);
        drho += dot(vel_i - wv(kk), gw);
    }

    // The reference's own walk (sim_solve.wgsl:436-466): lane l of the
    // particle's LANES = 8 takes the stencil cells l, l + 8, l + 16, l + 24
    // (cell numbering x fastest, sim_solve.wgsl:437-438) and sums the pairs
    // it accepts in its own walk order, lane 0 starting from the self term;
    // the eight partials fold with the xor butterfly. The list append is a
    // workgroup atomic (sim_solve.wgsl:463); the order taken is that of the
    // lanes in lockstep: for the lanes' k-th cell, their candidates' i-th
    // iterations in lane order (the atomics of one instruction resolve in
    // lane order), then the next i, then the next cell.
    void density_walk_faithful(const StepT<T>& s) {
        uint64_t over = 0;
#pragma omp parallel for schedule(static) reduction(+ : over)
        for (int64_t pi = 0; pi < (int64_t)n; ++pi) {
            uint32_t pidx = (uint32_t)pi;
            V3<T> pos_i = wp(pidx), vel_i = wv(pidx);
            T w_l[kRefLanes] = {}, rn_l[kRefLanes] = {}, gsq_l[kRefLanes] = {}, drho_l[kRefLanes] = {};
            V3<T> gsum_l[kRefLanes];
            for (uint32_t l = 0; l < kRefLanes; ++l) gsum_l[l] = v3(T(0), T(0), T(0));
            w_l[0] = k.sigma;
            uint32_t found = 0;
            int32_t base[3];
            cell_base(pidx, base);
            for (uint32_t round = 0; round * kRefLanes < 27; ++round) {
                uint32_t k0[kRefLanes], len[kRefLanes], longest = 0;
                for (uint32_t l = 0; l < kRefLanes; ++l) {
                    const int cell = (int)(round * kRefLanes + l);
                    k0[l] = len[l] = 0;
                    if (cell >= 27) continue;
                    int32_t cx = base[0] + (cell % 3) - 1;
                    int32_t cy = base[1] + (cell / 3) % 3 - 1;
                    int32_t cz = base[2] + cell / 9 - 1;
                    if (cx < 0 || cy < 0 || cz < 0 || cx >= (int32_t)p.dims[0] || cy >= (int32_t)p.dims[1] ||
                        cz >= (int32_t)p.dims[2])
                        continue;
                    uint32_t c = ((uint32_t)cz * p.dims[1] + (uint32_t)cy) * p.dims[0] + (uint32_t)cx;
                    k0[l] = starts[c];
                    len[l] = starts[c + 1] - starts[c];
                    longest = std::max(longest, len[l]);
                }
                for (uint32_t it = 0; it < longest; ++it) {
                    for (uint32_t l = 0; l < kRefLanes; ++l) {
                        if (it >= len[l]) continue;
                        const uint32_t kk = k0[l] + it;
                        if (kk == pidx) continue;
                        V3<T> x = pos_i - wp(kk);
                        T r2 = dot(x, x);
                        if (r2 >= k.support_sq) continue;
                        density_pair(vel_i, kk, x, r2, w_l[l], rn_l[l], gsq_l[l], drho_l[l], gsum_l[l]);
                        if (found < cap) nbr[(size_t)pidx * cap + found] = kk;
                        found++;
                    }
                }
            }
            nbr_n[pidx] = std::min(found, cap);
            if (found > cap) over += found - cap;
            density_finish(s, pidx, pos_i, vel_i, lane_fold(w_l, kRefLanes), lane_fold(rn_l, kRefLanes),
                           lane_fold(gsq_l, kRefLanes), lane_fold(drho_l, kRefLanes), lane_fold(gsum_l, kRefLanes));
        }
        nbr_over += over;
    }

    // Lane 0's end of density_div (sim_solve.wgsl:478-517): the mass, the
    // wall fill and gradient, rho, alpha and the divergence kappa / rho.
    void density_finish(const StepT<T>& s, uint32_t pidx, V3<T> pos_i, V3<T> vel_i, T w_sum, T rn, T grad_sq,
                        T drho, V3<T> grad_sum) {
        T rho = k.mass * w_sum;
        grad_sq = k.mass * k.mass * grad_sq;
        drho = k.mass * drho;
        grad_sum = k.mass * grad_sum;
        V3<T> lo = v3(k.wall_lo[0], k.wall_lo[1], k.wall_lo[2]);
        V3<T> hi = v3(k.wall_hi[0], k.wall_hi[1], k.wall_hi[2]);
        V3<T> tl = k.inv_h * (pos_i - lo);
        V3<T> th = k.inv_h * (hi - pos_i);
        T fill = 0;
        fill += wall_density(tl.x) + wall_density(th.x);
        fill += wall_density(tl.y) + wall_density(th.y);
        fill += wall_density(tl.z) + wall_density(th.z);
        fill -= wedge_sum(tl, th);
        rho += k.rho0 * fill;
        V3<T> wg = wall_grad_sum(k, pos_i);
        wgrad[4 * pidx] = wg.x;
        wgrad[4 * pidx + 1] = wg.y;
        wgrad[4 * pidx + 2] = wg.z;
        wgrad[4 * pidx + 3] = 0;
        grad_sum = grad_sum + wg;
        density[pidx] = rho;
        rho_near[pidx] = rn;
        T a = rho / tmax(dot(grad_sum, grad_sum) + grad_sq, K(1e-4));
        alpha[pidx] = a;
        drho += dot(vel_i, wg);
        kd[pidx] = tmax(drho, T(0)) * a / tmax(s.dt, K(1e-9)) / rho;
        if (err_on) err[pidx] = fixed_err(valf(tmax(drho, T(0)) * s.dt / k.rho0));
    }

    // Gradient of the pair (i, j), recomputed exactly as density_div did.
    inline V3<T> pair_grad(V3<T> pos_i, uint32_t j) const {
        V3<T> x = pos_i - wp(j);
        T r = tsqrt(dot(x, x));
        return grad_q(k, x, r, r * k.inv_h);
    }

    // kappa_partial + apply_kappa_wide, sim_solve.wgsl:596-621
    void kappa_apply(const StepT<T>& s) {
        std::vector<T> out((size_t)n * 3);
#pragma omp parallel for schedule(static)
        for (int64_t pi = 0; pi < (int64_t)n; ++pi) {
            uint32_t i = (uint32_t)pi;
            V3<T> pos_i = wp(i);
            T kd_i = kd[i];
            const uint32_t L = lanes();
            V3<T> part[kRefLanes];
            for (auto& q : part) q = v3(T(0), T(0), T(0));
            for (uint32_t m = 0; m < nbr_n[i]; ++m) {
                uint32_t j = nbr[(size_t)i * cap + m];
                part[m % L] = part[m % L] + (kd_i + kd[j]) * pair_grad(pos_i, j);
            }
            const V3<T> dv = lane_fold(part, L);
            V3<T> wg = v3(wgrad[4 * i], wgrad[4 * i + 1], wgrad[4 * i + 2]);
            V3<T> full = k.mass * dv + kd_i * wg;
            V3<T> v = wv(i) - s.dt * full;
            out[3 * i] = v.x;
            out[3 * i + 1] = v.y;
            out[3 * i + 2] = v.z;
        }
        for (uint32_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) wvel[4 * i + c] = out[3 * i + c];
    }

    // sim_solve.wgsl:524-592
    void forces_eval() {
#pragma omp parallel for schedule(static)
        for (int64_t pi = 0; pi < (int64_t)n; ++pi) {
            uint32_t i = (uint32_t)pi;
            const uint32_t L = lanes();
            V3<T> visc_l[kRefLanes], blend_l[kRefLanes], near_l[kRefLanes], tension_l[kRefLanes];
            T heat_l[kRefLanes] = {};
            for (uint32_t l = 0; l < kRefLanes; ++l)
                visc_l[l] = blend_l[l] = near_l[l] = tension_l[l] = v3(T(0), T(0), T(0));
            V3<T> pos_i = wp(i), vel_i = wv(i);
            T w_i = rho_near[i], temp_i = wtemp[i], rho_i = density[i];
            for (uint32_t m = 0; m < nbr_n[i]; ++m) {
                uint32_t j = nbr[(size_t)i * cap + m];
                const uint32_t l = m % L;
                V3<T>& visc = visc_l[l];
                V3<T>& blend = blend_l[l];
                V3<T>& nearf = near_l[l];
                V3<T>& tension = tension_l[l];
                T& heat = heat_l[l];
                V3<T> x = pos_i - wp(j);
                T r2 = dot(x, x);
                T r = tsqrt(r2);
                V3<T> g = grad_q(k, x, r, r * k.inv_h);
                T rho_j = density[j];
                T f = dot(x, g) / (r2 + k.eps_h2);
                T pair = k.mass / (rho_i * rho_j) * f;
                V3<T> dv = vel_i - wv(j);
                T wn = tmax(T(1) - r * k.inv_h, T(0));
                // x / max(r, 1e-5) and x / max(r, 1e-9) as x times one
                // reciprocal of r, shared when r >= 1e-5 (then both maxima
                // are r).
                T inv_rg = T(1) / tmax(r, K(1e-9));
                T inv_rm = r >= K(1e-5) ? inv_rg : T(1) / K(1e-5);
                T sn = (w_i + rho_near[j]) * wn * wn;
                if (faithful) {  // the shader's quotients, sim_solve.wgsl:559
                    const T dm = tmax(r, K(1e-5));
                    nearf = nearf + v3(sn * (x.x / dm), sn * (x.y / dm), sn * (x.z / dm));
                } else {
                    nearf = nearf + v3(sn * (x.x * inv_rm), sn * (x.y * inv_rm), sn * (x.z * inv_rm));
                }
                visc = visc + (k.visc2 * pair) * dv;
                blend = blend - (k.mass / rho_j * kernel_q(k, r * k.inv_h)) * dv;
                heat += k.heat_cond * pair * (temp_i - wtemp[j]);
                heat -= k.heat_visc * pair * dot(dv, dv);
                T corr = T(2) * k.rho0 / (rho_i + rho_j);
                T cm = corr * k.mass * cohesion(r, k.coh_c, k.coh_k, k.coh_near);
                if (faithful) {  // sim_solve.wgsl:567-568
                    const T dg = tmax(r, K(1e-9));
                    tension = tension - v3(cm * x.x / dg, cm * x.y / dg, cm * x.z / dg);
                } else {
                    tension = tension - v3(cm * x.x * inv_rg, cm * x.y * inv_rg, cm * x.z * inv_rg);
                }
            }
            const V3<T> visc = lane_fold(visc_l, L), blend = lane_fold(blend_l, L), nearf = lane_fold(near_l, L),
                        tension = lane_fold(tension_l, L);
            const T heat = lane_fold(heat_l, L);
            V3<T> adh = (k.beta * k.rho0) * wall_adh_sum(k, pos_i);
            V3<T> acc = visc + k.k_near * nearf + k.gamma * tension + adh;
            accel[4 * i] = acc.x;
            accel[4 * i + 1] = acc.y;
            accel[4 * i + 2] = acc.z;
            accel[4 * i + 3] = heat;
            xsph[4 * i] = blend.x;
            xsph[4 * i + 1] = blend.y;
            xsph[4 * i + 2] = blend.z;
            xsph[4 * i + 3] = 0;
            T rho = density[i];
            T half = K(0.5) * wpp[i];
            kd[i] = half / (rho * rho);
            pressure[i] = half;
        }
    }

    // sim_solve.wgsl:630-663
    void forces_den_apply(const StepT<T>& s) {
        std::vector<T> out((size_t)n * 3);
#pragma omp parallel for schedule(static)
        for (int64_t pi = 0; pi < (int64_t)n; ++pi) {
            uint32_t i = (uint32_t)pi;
            V3<T> pos_i = wp(i);
            T kd_i = kd[i];
            const uint32_t L = lanes();
            V3<T> part[kRefLanes];
            for (auto& q : part) q = v3(T(0), T(0), T(0));
            for (uint32_t m = 0; m < nbr_n[i]; ++m) {
                uint32_t j = nbr[(size_t)i * cap + m];
                part[m % L] = part[m % L] + (kd_i + kd[j]) * pair_grad(pos_i, j);
            }
            const V3<T> dv = lane_fold(part, L);
            V3<T> a = v3(accel[4 * i], accel[4 * i + 1], accel[4 * i + 2]);
            V3<T> r = pos_i, v = wv(i);
            V3<T> fict = -cross(s.domega, r) - cross(s.omega, cross(s.omega, r)) - T(2) * cross(s.omega, v);
            V3<T> xs = v3(xsph[4 * i], xsph[4 * i + 1], xsph[4 * i + 2]);
            V3<T> forced = fingers(s, r, v + s.dt * (s.force + fict + a) + s.c_xsph * xs, faithful);
            wtemp[i] = wtemp[i] + s.dt * accel[4 * i + 3];
            V3<T> wg = v3(wgrad[4 * i], wgrad[4 * i + 1], wgrad[4 * i + 2]);
            V3<T> full = k.mass * dv + kd_i * wg;
            V3<T> nv = forced - s.dt * full;
            out[3 * i] = nv.x;
            out[3 * i + 1] = nv.y;
            out[3 * i + 2] = nv.z;
        }
        for (uint32_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) wvel[4 * i + c] = out[3 * i + c];
    }

    // sim_solve.wgsl:674-698; with div, tolerance mode's divergence kappa:
    // density_div's formula (sim_solve.wgsl:508-516) from the current
    // velocities, no pressure.
    void den_kappa(const StepT<T>& s, bool div = false) {
#pragma omp parallel for schedule(static)
        for (int64_t pi = 0; pi < (int64_t)n; ++pi) {
            uint32_t i = (uint32_t)pi;
            V3<T> pos_i = wp(i), vel_i = wv(i);
            const uint32_t L = lanes();
            T part[kRefLanes] = {};
            for (uint32_t m = 0; m < nbr_n[i]; ++m) {
                uint32_t j = nbr[(size_t)i * cap + m];
                part[m % L] += dot(vel_i - wv(j), pair_grad(pos_i, j));
            }
            T drho = lane_fold(part, L);
            T rho = density[i];
            V3<T> wg = v3(wgrad[4 * i], wgrad[4 * i + 1], wgrad[4 * i + 2]);
            drho = k.mass * drho + dot(vel_i, wg);
            if (div) {
                kd[i] = tmax(drho, T(0)) * alpha[i] / tmax(s.dt, K(1e-9)) / rho;
                err[i] = fixed_err(valf(tmax(drho, T(0)) * s.dt / k.rho0));
                continue;
            }
            T rho_star = rho + s.dt * drho;
            T excess = tmax(rho_star - k.rho0, T(0));
            T kap = excess * alpha[i] / (s.dt * s.dt);
            kd[i] = kap / rho;
            pressure[i] += kap * rho;
            if (err_on) err[i] = fixed_err(valf(excess / k.rho0));
        }
    }

    // sim_solve.wgsl:709-756
    void integrate(const StepT<T>& s) {
        uint64_t clamps = 0;
#pragma omp parallel for schedule(static) reduction(+ : clamps)
        for (int64_t pi = 0; pi < (int64_t)n; ++pi) {
            uint32_t i = (uint32_t)pi;
            V3<T> v = wv(i);
            T was_w = wvel[4 * i + 3];
            T speed = tsqrt(dot(v, v));
            T ceiling = density[i] < K(0.5) * k.rho0 ? T(3) * s.v_clamp : s.v_clamp;
            if (speed > ceiling) {
                v = (ceiling / speed) * v;
                clamps++;
            }
            V3<T> pv = v3(wpvel[4 * i], wpvel[4 * i + 1], wpvel[4 * i + 2]);
            V3<T> dvv = v - pv;
            T acc = tsqrt(dot(dvv, dvv)) / s.dt;
            V3<T> np = wp(i) + s.dt * v;
            T out[3];
            T in[3] = {np.x, np.y, np.z};
            T vv[3] = {v.x, v.y, v.z};
            for (int a = 0; a < 3; ++a) {
                T c = tmin(tmax(in[a], k.wall_lo[a]), k.wall_hi[a]);
                out[a] = c;
                if (c != in[a]) vv[a] = T(0);
            }
            pos[4 * i] = out[0];
            pos[4 * i + 1] = out[1];
            pos[4 * i + 2] = out[2];
            pos[4 * i + 3] = 0;
            vel[4 * i] = vv[0];
            vel[4 * i + 1] = vv[1];
            vel[4 * i + 2] = vv[2];
            vel[4 * i + 3] = mix(was_w, acc, s.chase);
            pvel[4 * i] = vv[0];
            pvel[4 * i + 1] = vv[1];
            pvel[4 * i + 2] = vv[2];
            pvel[4 * i + 3] = mix(wpvel[4 * i + 3], pressure[i], s.chase);
            T dp = pressure[i] - wpp[i];
            pp[i] = pressure[i];
            T t = wtemp[i];
            temp[i] = t + t * T(EXPANSION) / (density[i] * T(HEAT_CAPACITY)) * dp;
            ids[i] = wids[i];
        }
        clamp_count += clamps;
    }

    // One substep in the reference dispatch order, render.rs:3175-3203.
    void substep(const StepBlock& sb, uint32_t refine) {
        StepT<T> s(sb, phys);
        seed(s.force.x, dforce[0]);
        seed(s.force.y, dforce[1]);
        seed(s.force.z, dforce[2]);
        grid_build();
        const bool div_on = div_tol > 0.0f && div_max > 1u, den_on = tol > 0.0f;
        err.assign(div_on || den_on ? n : 0u, 0u);
        err_on = div_on;
        density_div(s);
        uint32_t div_passes = 1;
        bool done = !div_on || mean_err() <= (double)div_tol;
        kappa_apply(s);       // div_apply
        while (!done && div_passes < div_max) {
            den_kappa(s, true);
            done = mean_err() <= (double)div_tol;
            kappa_apply(s);   // div_apply
            ++div_passes;
        }
        forces_eval();
        forces_den_apply(s);
        err_on = den_on;
        uint32_t passes = 0;
        for (const uint32_t cap_r = den_on ? max_passes : refine; passes < cap_r;) {
            den_kappa(s);
            const bool met = den_on && mean_err() <= (double)tol && passes + 1 >= min_passes;
            kappa_apply(s);   // den_apply
            ++passes;
            if (met) break;
        }
        err_on = false;
        last_passes[0] = passes;
        last_passes[1] = div_passes;
        if (div_on || den_on) {
            pass_hist[std::min<uint32_t>(passes, 256)] += 1;
            pass_hist[257 + std::min<uint32_t>(div_passes, 256)] += 1;
        }
        integrate(s);
    }

    // The zero-substep path of the reference's test driver (render.rs:3902-3937):
    // rebuild plus the density sweep alone.
    void density_only(const StepBlock& sb) {
        StepT<T> s(sb, phys);
        grid_build();
        density_div(s);
    }

    // sim_solve.wgsl:800-874 — sums in slot order, float accumulators.
    void reduce_stats() {
        float rho0 = p.rho0;
        float sum = 0, m = 0, lo = HUGE_F, hi = -HUGE_F;
        for (uint32_t i = 0; i < n; ++i) {
            float d = (float)density[i];
            float c = d / rho0 - 1.0f;
            sum += std::max(c, 0.0f);
            m = std::max(m, c);
            lo = std::min(lo, d);
            hi = std::max(hi, d);
        }
        if (faithful) {
            // The shader's order (sim_solve.wgsl:807-815, 767-778): thread l
            // of 256 sums slots l, l + 256, ..., then a tree halves the
            // partials (l += l + off for off = 128, 64, ..., 1).
            float part[256] = {};
            for (uint32_t l = 0; l < 256; ++l)
                for (uint32_t i = l; i < n; i += 256) part[l] += std::max((float)density[i] / rho0 - 1.0f, 0.0f);
            for (uint32_t off = 128; off > 0; off >>= 1)
                for (uint32_t l = 0; l < off; ++l) part[l] += part[l + off];
            sum = part[0];
        }
        stats[0] = sum / (float)n;
        stats[1] = std::max(m, 0.0f);
        stats[2] = lo;
        stats[3] = hi;
        float v_lo = HUGE_F, v_hi = -HUGE_F, a_lo = HUGE_F, a_hi = -HUGE_F;
        for (uint32_t i = 0; i < n; ++i) {
            float x = (float)vel[4 * i], y = (float)vel[4 * i + 1], z = (float)vel[4 * i + 2];
            float sp = std::sqrt(x * x + y * y + z * z);
            v_lo = std::min(v_lo, sp);
            v_hi = std::max(v_hi, sp);
            a_lo = std::min(a_lo, (float)vel[4 * i + 3]);
            a_hi = std::max(a_hi, (float)vel[4 * i + 3]);
        }
        stats[11] = v_lo;
        stats[6] = v_hi;
        stats[12] = a_lo;
        stats[13] = a_hi;
        float p_lo = HUGE_F, p_hi = -HUGE_F, t_lo = HUGE_F, t_hi = -HUGE_F;
        for (uint32_t i = 0; i < n; ++i) {
            float pw = (float)pvel[4 * i + 3];
            p_lo = std::min(p_lo, pw);
            p_hi = std::max(p_hi, pw);
            t_lo = std::min(t_lo, (float)temp[i]);
            t_hi = std::max(t_hi, (float)temp[i]);
        }
        stats[4] = p_lo;
        stats[5] = p_hi;
        stats[7] = t_lo;
        stats[8] = t_hi;
        stats[9] = (float)(uint32_t)clamp_count;
        stats[10] = (float)(uint32_t)nbr_over;
    }

    // ---- tracers, sim_tracers.wgsl:60-183 (float only) -------------------
    float ext(int a) const { return -(p.box_min[a] + p.cell); }
    void store_tracer(uint32_t* tr, uint32_t i, float x, float y, float z, float sp) const {
        float ex = ext(0), ey = ext(1);
        uint32_t ux = unorm16(x / ex * 0.5f + 0.5f), uy = unorm16(y / ey * 0.5f + 0.5f);
        tr[2 * i] = ux | (uy << 16);
        tr[2 * i + 1] = (uint32_t)f32_to_f16_rne(z) | ((uint32_t)f32_to_f16_rne(sp) << 16);
    }
    void respawn(uint32_t* tr, uint32_t slot, uint32_t r) const {
        uint32_t r1 = pcg(r), r2 = pcg(r1), r3 = pcg(r2), r4 = pcg(r3);
        float half_cell = p.cell * 0.5f;
        float jx = ((float)(r2 >> 20) / 4096.0f - 0.5f) * half_cell;
        float jy = ((float)(r3 >> 20) / 4096.0f - 0.5f) * half_cell;
        float jz = ((float)(r4 >> 20) / 4096.0f - 0.5f) * half_cell;
        uint32_t j = r1 % n;
        float e[3] = {ext(0), ext(1), ext(2)};
        float q[3] = {(float)pos[4 * j] + jx, (float)pos[4 * j + 1] + jy, (float)pos[4 * j + 2] + jz};
        for (int a = 0; a < 3; ++a) q[a] = std::min(std::max(q[a], -e[a]), e[a]);
        float vx = (float)vel[4 * j], vy = (float)vel[4 * j + 1], vz = (float)vel[4 * j + 2];
        store_tracer(tr, slot, q[0], q[1], q[2], std::sqrt(vx * vx + vy * vy + vz * vz));
    }
    void tracers_frame(const StepBlock& sb, uint32_t* tr, uint32_t ntr) {
        // splat, sim_tracers.wgsl:98-114 (integer sums: order-free)
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t c = cell_of(p, (float)pos[4 * i], (float)pos[4 * i + 1], (float)pos[4 * i + 2]) * 4u;
            for (int a = 0; a < 3; ++a)
                vel_grid[c + a] = (int32_t)((uint32_t)vel_grid[c + a] + (uint32_t)sat_i32((float)vel[4 * i + a] * FIXED));
            vel_grid[c + 3] = (int32_t)((uint32_t)vel_grid[c + 3] + 1u);
        }
        // resolve, sim_tracers.wgsl:122-138
        for (size_t c = 0; c < vel_grid.size(); ++c) {
            vel_flat[c] = vel_grid[c];
            vel_grid[c] = 0;
        }
        if (n == 0) return;
        // advect, sim_tracers.wgsl:140-183
        float dt = sb.dt;
        float decay = std::exp(-dt / T_CHARGE);
        float ex = ext(0), ey = ext(1), ez = ext(2);
        for (uint32_t i = 0; i < ntr; ++i) {
            uint32_t r0 = pcg(i ^ pcg(sb.seed));
            if ((float)r0 / 4294967295.0f < dt / TAU) {
                respawn(tr, i, r0);
                continue;
            }
            uint32_t t0 = tr[2 * i], t1 = tr[2 * i + 1];
            float px = ((float)(t0 & 0xffffu) / 65535.0f * 2.0f - 1.0f) * ex;
            float py = ((float)(t0 >> 16) / 65535.0f * 2.0f - 1.0f) * ey;
            float pz = f16_to_f32((uint16_t)(t1 & 0xffffu));
            float ow = f16_to_f32((uint16_t)(t1 >> 16));
            float pp3[3] = {px, py, pz};
            float gp[3], f[3];
            int32_t base[3];
            for (int a = 0; a < 3; ++a) {
                gp[a] = (pp3[a] - p.box_min[a]) / p.cell - 0.5f;
                int32_t b = sat_i32(std::floor(gp[a]));
                b = std::min(std::max(b, 0), (int32_t)p.dims[a] - 2);
                base[a] = b;
                f[a] = std::min(std::max(gp[a] - (float)(uint32_t)b, 0.0f), 1.0f);
            }
            float mx = 0, my = 0, mz = 0, wc = 0;
            for (uint32_t kk = 0; kk < 8; ++kk) {
                uint32_t o[3] = {kk & 1u, (kk >> 1) & 1u, (kk >> 2) & 1u};
                float w[3];
                for (int a = 0; a < 3; ++a) w[a] = mix(1.0f - f[a], f[a], (float)o[a]);
                float weight = w[0] * w[1] * w[2];
                uint32_t c = (((uint32_t)base[2] + o[2]) * p.dims[1] + ((uint32_t)base[1] + o[1])) * p.dims[0] +
                             ((uint32_t)base[0] + o[0]);
                const int32_t* cv = &vel_flat[(size_t)c * 4];
                mx += weight * (float)cv[0];
                my += weight * (float)cv[1];
                mz += weight * (float)cv[2];
                wc += weight * (float)cv[3];
            }
            float vx = 0, vy = 0, vz = 0;
            if (wc > 0.0f) {
                float den = wc * FIXED;
                vx = mx / den;
                vy = my / den;
                vz = mz / den;
            } else if ((float)pcg(r0 ^ 0x85ebca6bu) / 4294967295.0f < dt / TAU_STRAY) {
                respawn(tr, i, pcg(r0 ^ 0x632be59bu));
                continue;
            }
            float q[3] = {px + vx * dt, py + vy * dt, pz + vz * dt};
            float e[3] = {ex, ey, ez};
            for (int a = 0; a < 3; ++a) q[a] = std::min(std::max(q[a], -e[a]), e[a]);
            float sp = std::max(std::sqrt(vx * vx + vy * vy + vz * vz), ow * decay);
            store_tracer(tr, i, q[0], q[1], q[2], sp);
        }
    }
};

template <typename T>
void put(std::vector<T>& dst, const float* src, size_t count) {
    for (size_t i = 0; i < count; ++i) dst[i] = (T)src[i];
}
template <typename T>
void get(const std::vector<T>& src, float* dst, size_t count) {
    for (size_t i = 0; i < count; ++i) dst[i] = (float)src[i];
}
template <typename T>
void getd(const std::vector<T>& src, double* dst, size_t count) {
    for (size_t i = 0; i < count; ++i) dst[i] = (double)src[i];
}

struct Handle {
    int precision;  // 32, 64, 128 (the dual-number FP64 run) or 96 (the dual-number FP32 run)
    Sim<float>* f;
    Sim<double>* d;
    Sim<Dual>* a;
    Sim<DualF>* af;
};

template <typename F>
auto dispatch(Handle* h, F&& fn) {
    if (h->precision == 128) return fn(*h->a);
    if (h->precision == 96) return fn(*h->af);
    if (h->precision == 64) return fn(*h->d);
    return fn(*h->f);
}

}  // namespace

extern "C" {

// Creates an oracle sim from the 64-byte SimParams block; precision 32, 64,
// 128 (dual numbers over double) or 96 (dual numbers over float values).
void* oracle_create(const void* params64, uint32_t nbr_cap, int precision) {
    Params p;
    std::memcpy(&p, params64, 64);
    Handle* h = new Handle{precision, nullptr, nullptr, nullptr, nullptr};
    if (precision == 128)
        h->a = new Sim<Dual>(p, nbr_cap);
    else if (precision == 96)
        h->af = new Sim<DualF>(p, nbr_cap);
    else if (precision == 64)
        h->d = new Sim<double>(p, nbr_cap);
    else
        h->f = new Sim<float>(p, nbr_cap);
    return h;
}

void oracle_destroy(void* hv) {
    Handle* h = (Handle*)hv;
    delete h->f;
    delete h->d;
    delete h->a;
    delete h->af;
    delete h;
}

// {mu, k_near, tension_scale, adhesion_scale, xsph_rate}
void oracle_set_physics(void* hv, const float* phys5) {
    Physics ph;
    ph.mu = phys5[0];
    ph.k_near = phys5[1];
    ph.tension_scale = phys5[2];
    ph.adhesion_scale = phys5[3];
    ph.xsph_rate = phys5[4];
    dispatch((Handle*)hv, [&](auto& s) -> int {
        s.set_physics(ph);
        return 0;
    });
}

// Tolerance mode (see Sim::tol): a tolerance <= 0 keeps that solve's
// reference schedule.
void oracle_set_tolerance(void* hv, float tol, uint32_t min_passes, uint32_t max_passes, float div_tol,
                          uint32_t div_max) {
    dispatch((Handle*)hv, [&](auto& s) {
        s.tol = tol > 0.0f ? tol : 0.0f;
        s.min_passes = min_passes > 0 ? min_passes : 1u;
        s.max_passes = max_passes;
        s.div_tol = div_tol > 0.0f ? div_tol : 0.0f;
        s.div_max = div_tol > 0.0f ? div_max : 1u;
        return 0;
    });
}

// The last substep's passes {density, divergence} and the pass histograms
// (2 x 257 counts, as the HIP step's).
void oracle_passes(void* hv, uint32_t* last2, uint64_t* hist514) {
    dispatch((Handle*)hv, [&](auto& s) {
        if (last2) {
            last2[0] = s.last_passes[0];
            last2[1] = s.last_passes[1];
        }
        if (hist514)
            for (int i = 0; i < 2 * 257; ++i) hist514[i] = s.pass_hist[i];
        return 0;
    });
}

// The line-faithful mode (1) or the kernels' order (0, the default); see
// Sim::faithful.
void oracle_set_faithful(void* hv, int on) {
    dispatch((Handle*)hv, [&](auto& s) {
        s.faithful = on != 0;
        return 0;
    });
}

// Dual runs (precision 128 or 96): tangents of the resting set's positions (0,
// n x 4), velocities (1, n x 4) or previous pressures (2, n).
void oracle_set_tangent(void* hv, int which, const double* d) {
    dispatch((Handle*)hv, [&](auto& s) {
        auto* v = which == 0 ? &s.pos : which == 1 ? &s.vel : &s.pp;
        const size_t cnt = which == 2 ? s.n : (size_t)s.n * 4;
        for (size_t i = 0; i < cnt; ++i) seed((*v)[i], d[i]);
        return 0;
    });
}

void oracle_get_tangent(void* hv, int which, double* out) {
    dispatch((Handle*)hv, [&](auto& s) {
        const auto* v = which == 0 ? &s.pos : which == 1 ? &s.vel : &s.pp;
        const size_t cnt = which == 2 ? s.n : (size_t)s.n * 4;
        for (size_t i = 0; i < cnt; ++i) out[i] = tangent((*v)[i]);
        return 0;
    });
}

// The tangent of the next substeps' gravity (kept until changed).
void oracle_seed_force(void* hv, const double* d3) {
    dispatch((Handle*)hv, [&](auto& s) {
        for (int a = 0; a < 3; ++a) s.dforce[a] = d3[a];
        return 0;
    });
}

// The tangent of the viscosity (after oracle_set_physics).
void oracle_seed_mu(void* hv, double dmu) {
    dispatch((Handle*)hv, [&](auto& s) {
        s.seed_mu(dmu);
        return 0;
    });
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Resting-set upload; any pointer may be null (left unchanged).
void oracle_set_state(void* hv, const float* pos4, const float* vel4, const float* prev_vel4,
                      const float* prev_pressure, const float* temperature) {
    dispatch((Handle*)hv, [&](auto& s) {
        size_t n = s.n;
        if (pos4) put(s.pos, pos4, n * 4);
        if (vel4) put(s.vel, vel4, n * 4);
        if (prev_vel4) put(s.pvel, prev_vel4, n * 4);
        if (prev_pressure) put(s.pp, prev_pressure, n);
        if (temperature) put(s.temp, temperature, n);
        return 0;
    });
}

void oracle_get_state(void* hv, float* pos4, float* vel4, float* prev_vel4, float* prev_pressure,
                      float* temperature, uint32_t* ids) {
    dispatch((Handle*)hv, [&](auto& s) {
        size_t n = s.n;
        if (pos4) get(s.pos, pos4, n * 4);
        if (vel4) get(s.vel, vel4, n * 4);
        if (prev_vel4) get(s.pvel, prev_vel4, n * 4);
        if (prev_pressure) get(s.pp, prev_pressure, n);
        if (temperature) get(s.temp, temperature, n);
        if (ids) std::memcpy(ids, s.ids.data(), n * 4);
        return 0;
    });
}

void oracle_get_state_f64(void* hv, double* pos4, double* vel4, double* prev_vel4, double* prev_pressure,
                          double* temperature) {
    dispatch((Handle*)hv, [&](auto& s) {
        size_t n = s.n;
        if (pos4) getd(s.pos, pos4, n * 4);
        if (vel4) getd(s.vel, vel4, n * 4);
        if (prev_vel4) getd(s.pvel, prev_vel4, n * 4);
        if (prev_pressure) getd(s.pp, prev_pressure, n);
        if (temperature) getd(s.temp, temperature, n);
        return 0;
    });
}

// Scratch fields by name: 0 density, 1 alpha, 2 kd, 3 pressure, 4 rho_near,
// 5 accel (x4), 6 xsph (x4), 7 wall_grad (x4), 8 working pos (x4),
// 9 working vel (x4). Returns the float count written.
size_t oracle_get_field(void* hv, int which, double* out) {
    return dispatch((Handle*)hv, [&](auto& s) -> size_t {
        const auto* v = &s.density;
        size_t per = 1;
        switch (which) {
            case 0: v = &s.density; break;
            case 1: v = &s.alpha; break;
            case 2: v = &s.kd; break;
            case 3: v = &s.pressure; break;
            case 4: v = &s.rho_near; break;
            case 5: v = &s.accel; per = 4; break;
            case 6: v = &s.xsph; per = 4; break;
            case 7: v = &s.wgrad; per = 4; break;
            case 8: v = &s.wpos; per = 4; break;
            case 9: v = &s.wvel; per = 4; break;
            default: return 0;
        }
        size_t cnt = (size_t)s.n * per;
        getd(*v, out, cnt);
        return cnt;
    });
}

// Grid of the last rebuild: cell id per resting particle, per-cell counts,
// starts (cells + 1), slot -> resting-index permutation, neighbour counts.
void oracle_get_grid(void* hv, uint32_t* cellid, uint32_t* counts, uint32_t* starts, uint32_t* perm,
                     uint32_t* nbr_n) {
    dispatch((Handle*)hv, [&](auto& s) {
        if (cellid) std::memcpy(cellid, s.cellid.data(), (size_t)s.n * 4);
        if (counts) std::memcpy(counts, s.hist.data(), (size_t)s.cells * 4);
        if (starts) std::memcpy(starts, s.starts.data(), ((size_t)s.cells + 1) * 4);
        if (perm) std::memcpy(perm, s.perm.data(), (size_t)s.n * 4);
        if (nbr_n) std::memcpy(nbr_n, s.nbr_n.data(), (size_t)s.n * 4);
        return 0;
    });
}

void oracle_get_nbr(void* hv, uint32_t* nbr) {
    dispatch((Handle*)hv, [&](auto& s) {
        std::memcpy(nbr, s.nbr.data(), s.nbr.size() * 4);
        return 0;
    });
}

void oracle_substep(void* hv, const void* step144, uint32_t refine) {
    StepBlock sb;
    std::memcpy(&sb, step144, sizeof sb);
    dispatch((Handle*)hv, [&](auto& s) {
        s.substep(sb, refine);
        return 0;
    });
}

void oracle_density_only(void* hv, const void* step144) {
    StepBlock sb;
    std::memcpy(&sb, step144, sizeof sb);
    dispatch((Handle*)hv, [&](auto& s) {
        s.density_only(sb);
        return 0;
    });
}

// One kernel of the substep by name (grid, density_div, div_apply,
// forces_eval, forces_den_apply, den_kappa, den_apply, integrate), so tests
// can check invariants of the pieces. Returns -1 for an unknown name.
int oracle_kernel(void* hv, const char* name, const void* step144) {
    StepBlock sb;
    std::memcpy(&sb, step144, sizeof sb);
    std::string k(name);
    return dispatch((Handle*)hv, [&](auto& s) -> int {
        using S = std::remove_reference_t<decltype(s)>;
        using T = std::remove_reference_t<decltype(s.density[0])>;
        (void)sizeof(S);
        StepT<T> st(sb, s.phys);
        if (k == "grid") s.grid_build();
        else if (k == "density_div") s.density_div(st);
        else if (k == "div_apply" || k == "den_apply") s.kappa_apply(st);
        else if (k == "forces_eval") s.forces_eval();
        else if (k == "forces_den_apply") s.forces_den_apply(st);
        else if (k == "den_kappa") s.den_kappa(st);
        else if (k == "div_kappa") {
            s.err.assign(s.n, 0u);
            s.den_kappa(st, true);
        }
        else if (k == "integrate") s.integrate(st);
        else return -1;
        return 0;
    });
}

void oracle_grid_build(void* hv) {
    dispatch((Handle*)hv, [&](auto& s) {
        s.grid_build();
        return 0;
    });
}

void oracle_reduce_stats(void* hv, float* out14) {
    dispatch((Handle*)hv, [&](auto& s) {
        s.reduce_stats();
        std::memcpy(out14, s.stats, sizeof s.stats);
        return 0;
    });
}

void oracle_counters(void* hv, uint64_t* clamp_count, uint64_t* nbr_over) {
    dispatch((Handle*)hv, [&](auto& s) {
        *clamp_count = s.clamp_count;
        *nbr_over = s.nbr_over;
        return 0;
    });
}

// splat + resolve + advect for `ntr` packed tracers (in place), float only.
int oracle_tracers(void* hv, const void* step144, uint32_t* tracers, uint32_t ntr) {
    Handle* h = (Handle*)hv;
    if (h->precision != 32) return -1;
    StepBlock sb;
    std::memcpy(&sb, step144, sizeof sb);
    h->f->tracers_frame(sb, tracers, ntr);
    return 0;
}

// Standalone pieces the tests pin against closed forms.
uint32_t oracle_cell_of(const void* params64, float x, float y, float z) {
    Params p;
    std::memcpy(&p, params64, 64);
    return cell_of(p, x, y, z);
}
double oracle_wall_density(double t) { return wall_density<double>(t); }
double oracle_wall_gradient(double t) { return wall_gradient<double>(t); }
double oracle_wedge(double a, double b) { return wedge<double>(a, b); }
double oracle_wedge_d(double a, double b) { return wedge_d<double>(a, b); }
double oracle_wall_adhesion(double t) { return wall_adhesion<double>(t); }
float oracle_exp_fixed(float x) { return exp_fixed(x); }
float oracle_wall_density_f32(float t) { return wall_density<float>(t); }
uint32_t oracle_pcg(uint32_t x) { return pcg(x); }
uint16_t oracle_f16(float x) { return f32_to_f16_rne(x); }
float oracle_f16_to_f32(uint16_t x) { return f16_to_f32(x); }

}  // extern "C"
