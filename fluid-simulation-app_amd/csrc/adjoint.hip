// The adjoint of the DFSPH substep for gfx950: reverse sweeps that carry the
// gradient of a loss on the final state back through one substep.
//
// What is differentiated (the whole step at any physics setting, with no
// touches; the grid permutation and neighbour lists held fixed per substep,
// as recorded by the taped forward):
//   density_div   rho, rho_near, alpha and the divergence kappa
//                 (sim_solve.wgsl:411-518), including the wall fill and wall
//                 gradient (:212-328)
//   div_apply     the divergence-free correction (:594-621, 665-668)
//   forces_eval   Morris viscosity, near pressure, cohesion with its
//                 density correction, wall adhesion and the XSPH blend
//                 (:524-583), and the warm start (:586-590)
//   forces_den_apply  gravity, the rotating frame's fictitious forces, the
//                 accelerations and the XSPH blend, then the warm-start
//                 correction (:630-663)
//   den_kappa / den_apply   the constant-density Jacobi passes (:670-703),
//                 reversed pass by pass: each transposed Jacobi pass is a
//                 neighbour sweep of the same shape as the forward one
//   integrate     the speed clamp and the wall clamp (:705-756)
// Temperature, the lens EMAs and the stats do not feed back into the motion
// and carry no adjoint; the finger pull (touches) has no reverse, and the
// forward refuses substeps that have touches.
//
// Pair geometry: r = x_i - x_j, rho = |r|, grad W_ij = g r with g the list's
// dW/r, and the Hessian H_ij = g I + hh r r^T with hh = (dw'(q)/h - g) / rho^2.
// The lists are symmetric (the squared distance is bit-identical both ways),
// so every pair term's adjoint is gathered by both particles from their own
// lists. The sweeps run four lanes a particle as the forward does (lane l
// takes list entries l, l + 4, ...; the partials fold across the quad), the
// per-particle work one thread a particle.
#include "fluid_common.hpp"
#include "fluid_kernels.hpp"
#include "adjoint.hpp"

namespace fluid {

namespace {

constexpr uint32_t kAdjBlock = 256;
constexpr uint32_t kQuadBlock = 256 * kListLanes;  // 256 particles a workgroup

// A particle's lane in the four-lane sweeps, with the particle's list count
// and where the lane's entries past the list's fixed rows start.
struct Quad {
    uint32_t p, lane, gid, nn, tail;
    bool live;
};
// Every lane of the wave calls it (the tail start is a sum over the wave).
__device__ __forceinline__ Quad quad(uint32_t n, const SolveBufs& b) {
    Quad q;
    q.gid = blockIdx.x * kQuadBlock + threadIdx.x;
    q.p = q.gid / kListLanes;
    q.lane = q.gid % kListLanes;
    q.live = q.p < n;
    q.nn = q.live ? b.nbr_n[q.p] : 0u;
    const uint32_t len = list_tail_len(q.nn);
    const uint32_t off = wave_inclusive_sum((__lane_id() & (kListLanes - 1)) == 0 ? len : 0u) - len;
    q.tail = list_tail(b.stride, b.cap, q.p, off) + q.lane;
    return q;
}
// Calls f(neighbour slot, dW/r) for the lane's share of particle p's list
// (entries lane, lane + 4, ...: the fixed rows, then the wave's tail block).
// The taped forward stages nothing, so its lists hold slots.
template <class F>
__device__ __forceinline__ void for_lane(const SolveBufs& b, const Quad& q, F&& f) {
    const uint32_t step = b.stride * kListLanes;  // list indices fit 32 bits
    uint32_t t = 0;
    for (uint32_t m = q.lane; m < q.nn; m += kListLanes, ++t) {
        const uint32_t e = t < kListFixed ? t * step + q.gid : q.tail + (t - kListFixed) * kListLanes;
        f(b.nbr32[e], b.nbrg[e]);
    }
}

__device__ __forceinline__ f3 ld3(const float4* a, uint32_t i) { return xyz(a[i]); }
__device__ __forceinline__ void add3(float4* a, uint32_t i, f3 v) {
    float4 w = a[i];
    a[i] = make_float4(w.x + v.x, w.y + v.y, w.z + v.z, w.w);
}

struct PairGeom {
    f3 r;
    float r2, d, g, hh;  // |r|^2, |r|
};

// dw'(q): the q-derivative of the kernel-gradient magnitude of grad_scale
// (sim_solve.wgsl:184-196).
__device__ __forceinline__ PairGeom pair_geom(const SceneK& k, f3 xi, f3 xj, float g) {
    PairGeom P;
    P.r = sub(xi, xj);
    const float r2 = dot3(P.r, P.r);
    const float rho = sqrt_rn(r2);
    const float q = rho * k.inv_h;
    const float dwp = q < 1.0f ? k.sigma_h * (-3.0f + 4.5f * q) : k.sigma_h * (1.5f * (2.0f - q));
    P.r2 = r2;
    P.d = rho;
    P.g = g;
    P.hh = rho > 1e-9f ? (dwp * k.inv_h - g) / r2 : 0.0f;
    return P;
}
__device__ __forceinline__ f3 hess(const PairGeom& P, f3 u) { return add(scl(P.g, u), scl(P.hh * dot3(P.r, u), P.r)); }


// Transposed 3x3 wall-gradient Jacobian times u: J rows are d wg / dx_a
// stored as columns a (J[a] = d wg / d x_a), so J^T u = (J[0].u, J[1].u, J[2].u).
__device__ __forceinline__ f3 jt(const float4* J, uint32_t i, f3 u) {
    return mk(dot3(xyz(J[3 * i]), u), dot3(xyz(J[3 * i + 1]), u), dot3(xyz(J[3 * i + 2]), u));
}

// Per particle: the gradient of the wall fill (grad F) and the Jacobian of
// wall_grad_sum, by dual evaluation of the same polynomials.
__device__ __forceinline__ void wall_jacobians(const SceneK& k, uint32_t i, const float4* __restrict__ pos,
                                               float4* __restrict__ dfill, float4* __restrict__ J) {
    const f3 p = xyz(pos[i]);
    // Every wall term vanishes when all six wall distances are at least the
    // support (t >= 2): wall_density, wall_gradient, and the wedges (which
    // need two of them below 2).
    const float ih = fminf(k.inv_h, k.inv_h_wall);
    float dmin = fminf(fminf(p.x - k.wall_lo[0], k.wall_hi[0] - p.x), fminf(p.y - k.wall_lo[1], k.wall_hi[1] - p.y));
    dmin = fminf(dmin, fminf(p.z - k.wall_lo[2], k.wall_hi[2] - p.z));
    if (dmin * ih >= 2.0f) {
        dfill[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (int a = 0; a < 3; ++a) J[3 * i + a] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    float df[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const Dn px(p.x, a == 0 ? 1.0f : 0.0f), py(p.y, a == 1 ? 1.0f : 0.0f), pz(p.z, a == 2 ? 1.0f : 0.0f);
        df[a] = wall_fill_t(k, px, py, pz).d;
        Dn gx, gy, gz;
        wall_grad_sum_t(k, px, py, pz, gx, gy, gz);
        J[3 * i + a] = make_float4(gx.d, gy.d, gz.d, 0.0f);
    }
    dfill[i] = make_float4(df[0], df[1], df[2], 0.0f);
}

// integrate (sim_solve.wgsl:709-756) in reverse. lam_x, lam_v arrive as the
// adjoints of the resting position and velocity, which integrate writes at
// the working slot; lam_p is the adjoint of the pressure it stores as the
// next substep's previous pressure. Outputs, in place: lam_x of the working
// position and lam_v of the velocity entering integrate.
// Also, per particle: the wall Jacobians of the substep and zeroed
// lam_rho / lam_alpha accumulators.
__global__ __launch_bounds__(kAdjBlock) void k_rev_integrate(SceneK k, StepK s, uint32_t n,
                                                             const float4* __restrict__ pos,
                                                             const float4* __restrict__ vin,
                                                             const float* __restrict__ density,
                                                             float4* __restrict__ lam_x, float4* __restrict__ lam_v,
                                                             float4* __restrict__ dfill, float4* __restrict__ J,
                                                             float* __restrict__ lam_rho,
                                                             float* __restrict__ lam_alpha) {
    const uint32_t i = blockIdx.x * kAdjBlock + threadIdx.x;
    if (i >= n) return;
    lam_rho[i] = 0.0f;
    lam_alpha[i] = 0.0f;
    wall_jacobians(k, i, pos, dfill, J);
    f3 v = xyz(vin[i]);
    const float speed = sqrtf(dot3(v, v));
    const float rho = density[i];
    const float ceiling = rho < 0.5f * k.rho0 ? 3.0f * s.v_clamp : s.v_clamp;
    const bool clamped = speed > ceiling;
    if (clamped) v = scl(ceiling / speed, v);
    const f3 p = xyz(pos[i]);
    const float in[3] = {p.x + s.dt * v.x, p.y + s.dt * v.y, p.z + s.dt * v.z};
    const f3 lx = xyz(lam_x[i]), lvo = xyz(lam_v[i]);
    const float lxa[3] = {lx.x, lx.y, lx.z}, lva[3] = {lvo.x, lvo.y, lvo.z};
    float ox[3], ov[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float c = wmin(wmax(in[a], k.wall_lo[a]), k.wall_hi[a]);
        const bool hit = c != in[a];
        ox[a] = hit ? 0.0f : lxa[a];
        ov[a] = hit ? 0.0f : lva[a] + s.dt * lxa[a];
    }
    f3 lv = mk(ov[0], ov[1], ov[2]);
    if (clamped) {
        const f3 u = scl(1.0f / speed, xyz(vin[i]));
        lv = scl(ceiling / speed, sub(lv, scl(dot3(u, lv), u)));
    }
    lam_x[i] = make_float4(ox[0], ox[1], ox[2], 0.0f);
    lam_v[i] = make_float4(lv.x, lv.y, lv.z, 0.0f);
}

// An apply sweep in reverse (div_apply, the warm start of forces_den_apply,
// den_apply): v_out = v_in - dt (m sum (kd_i + kd_j) gradW_ij + kd_i wg_i).
// lam (the adjoint of v_out) passes to v_in unchanged; this sweep yields
// lam_kd and adds the position terms to lam_x.
// The per-particle reverse of den_kappa (below) that follows den_apply's.
__device__ __forceinline__ void den_kappa_pre(const StepK& s, uint32_t i, float lk, const float* __restrict__ rho,
                                              const float* __restrict__ alpha, const float* __restrict__ kd,
                                              const float* __restrict__ lam_p, float* __restrict__ lam_D,
                                              float* __restrict__ lam_rho, float* __restrict__ lam_alpha);

// With DEN, the reverse of den_kappa's per-particle part follows at once
// (its kd is this apply's kd), giving lam_D for the divergence sweep.
template <bool DEN>
__global__ __launch_bounds__(kQuadBlock) void k_rev_apply(SceneK k, StepK st, uint32_t n, SolveBufs b,
                                                          const float4* __restrict__ pos,
                                                          const float* __restrict__ kd,
                                                          const float4* __restrict__ wg, const float4* __restrict__ J,
                                                          const float4* __restrict__ lam, float* __restrict__ lam_kd,
                                                          float4* __restrict__ lam_x, const float* __restrict__ rho,
                                                          const float* __restrict__ alpha,
                                                          const float* __restrict__ lam_p, float* __restrict__ lam_D,
                                                          float* __restrict__ lam_rho, float* __restrict__ lam_alpha) {
    const float dt = st.dt;
    const Quad q = quad(n, b);
    float s_lk = 0.0f;
    f3 s_x = mk(0.0f, 0.0f, 0.0f), li = s_x;
    float ki = 0.0f;
    if (q.live) {
        const f3 xi = ld3(pos, q.p);
        li = ld3(lam, q.p);
        ki = kd[q.p];
        for_lane(b, q, [&](uint32_t j, float g) {
            const PairGeom P = pair_geom(k, xi, ld3(pos, j), g);
            const f3 dl = sub(li, ld3(lam, j));
            s_lk += dot3(dl, scl(P.g, P.r));
            s_x = add(s_x, scl(ki + kd[j], hess(P, dl)));
        });
    }
    s_lk = fold4(s_lk);
    s_x = fold4(s_x);
    if (!q.live || q.lane != 0) return;
    const uint32_t i = q.p;
    const float lk = -dt * (k.mass * s_lk + dot3(li, ld3(wg, i)));
    const f3 w = jt(J, i, li);
    add3(lam_x, i, scl(-dt, add(scl(k.mass, s_x), scl(ki, w))));
    if (DEN)
        den_kappa_pre(st, i, lk, rho, alpha, kd, lam_p, lam_D, lam_rho, lam_alpha);
    else
        lam_kd[i] = lk;
}

// den_kappa in reverse, per particle: from lam_kd of the pass (and the
// pressure adjoint lam_p, p += kappa rho) to lam_D, adding to lam_rho and
// lam_alpha. kappa = max(rho + dt D - rho0, 0) alpha / dt^2, kd = kappa / rho.
__device__ __forceinline__ void den_kappa_pre(const StepK& s, uint32_t i, float lk, const float* __restrict__ rho,
                                              const float* __restrict__ alpha, const float* __restrict__ kd,
                                              const float* __restrict__ lam_p, float* __restrict__ lam_D,
                                              float* __restrict__ lam_rho, float* __restrict__ lam_alpha) {
    const float r = rho[i], a = alpha[i], k = kd[i], lp = lam_p[i];
    const float kap = k * r;
    const float lkap = lk / r + lp * r;
    float lr = -lk * k / r + lp * kap;
    float lD = 0.0f;
    if (k > 0.0f) {
        const float dt2 = s.dt * s.dt;
        const float le = lkap * a / dt2;
        lam_alpha[i] += lkap * k * r / a;  // lkap * e / dt^2, e = kd rho dt^2 / alpha
        lr += le;
        lD = s.dt * le;
    }
    lam_rho[i] += lr;
    lam_D[i] = lD;
}

// The sweep of a divergence-type term D_i = m sum (v_i - v_j).gradW_ij +
// v_i.wg_i in reverse (den_kappa, and density_div's D): lam_v_i = m sum
// (lam_D_i + lam_D_j) gradW_ij + lam_D_i wg_i, and the position terms.
__global__ __launch_bounds__(kQuadBlock) void k_rev_divergence(SceneK k, uint32_t n, SolveBufs b,
                                                               const float4* __restrict__ pos,
                                                               const float4* __restrict__ vel,
                                                               const float4* __restrict__ wg,
                                                               const float4* __restrict__ J,
                                                               const float* __restrict__ lam_D,
                                                               float4* __restrict__ lam_v,
                                                               float4* __restrict__ lam_x) {
    const Quad q = quad(n, b);
    f3 s_v = mk(0.0f, 0.0f, 0.0f), s_x = s_v, vi = s_v;
    float di = 0.0f;
    if (q.live) {
        const f3 xi = ld3(pos, q.p);
        vi = ld3(vel, q.p);
        di = lam_D[q.p];
        for_lane(b, q, [&](uint32_t j, float g) {
            const PairGeom P = pair_geom(k, xi, ld3(pos, j), g);
            const float dj = lam_D[j];
            // D_j holds (v_j - v_i).gradW_ji = v_i.gradW_ij + ..., so both ends add
            s_v = add(s_v, scl(di + dj, scl(P.g, P.r)));
            s_x = add(s_x, scl(di + dj, hess(P, sub(vi, ld3(vel, j)))));
        });
    }
    s_v = fold4(s_v);
    s_x = fold4(s_x);
    if (!q.live || q.lane != 0) return;
    const uint32_t i = q.p;
    add3(lam_v, i, add(scl(k.mass, s_v), scl(di, ld3(wg, i))));
    add3(lam_x, i, add(scl(k.mass, s_x), scl(di, jt(J, i, vi))));
}

// forces_den_apply's first half in reverse: v2 = v1 + dt (g + f(x, v1) + a)
// + c_xsph xs with the rotating frame's fictitious force f = -dw x x -
// w x (w x x) - 2 w x v1. lam_f, the adjoint of v2, arrives in lam_v and
// leaves as v1's: lam_f + 2 dt w x lam_f; lam_x gains dt (dw x lam_f -
// w x (w x lam_f)); lam_a = dt lam_f, lam_xs = c_xsph lam_f; the gravity
// adjoint of the substep is dt sum lam_f (one atomic per block). The warm
// start kd_w = 0.5 pp / rho^2 and p = 0.5 pp: lam_pp and lam_rho.
__global__ __launch_bounds__(kAdjBlock) void k_rev_forces_den(StepK s, uint32_t n, const float* __restrict__ rho,
                                                              const float* __restrict__ pp,
                                                              float4* __restrict__ lam_v,
                                                              const float* __restrict__ lam_kdw,
                                                              const float* __restrict__ lam_p,
                                                              float4* __restrict__ lam_a, float4* __restrict__ lam_xs,
                                                              float4* __restrict__ lam_x, float* __restrict__ lam_pp,
                                                              float* __restrict__ lam_rho, float* __restrict__ lam_g) {
    __shared__ float red[3][kAdjBlock];
    const uint32_t i = blockIdx.x * kAdjBlock + threadIdx.x;
    f3 lf = mk(0.0f, 0.0f, 0.0f);
    if (i < n) {
        lf = ld3(lam_v, i);
        const f3 om = mk(s.omega[0], s.omega[1], s.omega[2]);
        const f3 dom = mk(s.domega[0], s.domega[1], s.domega[2]);
        const f3 lv = add(lf, scl(2.0f * s.dt, cross3(om, lf)));
        lam_v[i] = make_float4(lv.x, lv.y, lv.z, 0.0f);
        add3(lam_x, i, scl(s.dt, sub(cross3(dom, lf), cross3(om, cross3(om, lf)))));
        lam_a[i] = make_float4(s.dt * lf.x, s.dt * lf.y, s.dt * lf.z, 0.0f);
        lam_xs[i] = make_float4(s.c_xsph * lf.x, s.c_xsph * lf.y, s.c_xsph * lf.z, 0.0f);
        const float r = rho[i], lk = lam_kdw[i];
        lam_pp[i] = 0.5f * lk / (r * r) + 0.5f * lam_p[i];
        lam_rho[i] += -lk * pp[i] / (r * r * r);
    }
    red[0][threadIdx.x] = lf.x;
    red[1][threadIdx.x] = lf.y;
    red[2][threadIdx.x] = lf.z;
    __syncthreads();
    for (uint32_t o = kAdjBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 3) atomicAdd(&lam_g[threadIdx.x], s.dt * red[threadIdx.x][0]);
}

// d cohesion / dr (sim_solve.wgsl:201-210): a = (c - r)^3 r^3, a' = 3 (c -
// r)^2 r^2 (c - 2 r), doubled on the near branch.
__device__ __forceinline__ float cohesion_dr(const SceneK& k, float r) {
    const float c = k.coh_c;
    if (r >= c || r < 1e-9f) return 0.0f;
    const float cr = c - r;
    const float da = 3.0f * cr * cr * r * r * (c - 2.0f * r);
    return 2.0f * r > c ? k.coh_k * da : 2.0f * k.coh_k * da;
}

// d wall_adhesion / dt (sim_solve.wgsl:334-340): u = 2 - t where it lies in
// (0, 2), else the clamp holds u and the derivative is 0.
__device__ __forceinline__ float wall_adhesion_dt(float t) {
    const float u = 2.0f - t;
    if (!(u > 0.0f && u < 2.0f)) return 0.0f;
    const float P = 1.7847015e-3f + u * (3.3392196e-3f + u * (-4.3394677e-3f + u * (1.6711868e-3f + u * -2.1788750e-4f)));
    const float dP = 3.3392196e-3f + u * (2.0f * -4.3394677e-3f + u * (3.0f * 1.6711868e-3f + u * (4.0f * -2.1788750e-4f)));
    return -(2.0f * u * P + u * u * dP);
}

// forces_eval in reverse (sim_solve.wgsl:544-583). With la = lam_a and lxs =
// lam_xs, the pair terms of particle i and their adjoints, gathered by i
// over its (symmetric) list:
//   viscosity  a_i += visc2 P_ij (v_i - v_j), P_ij = m f(|r|) / (rho_i rho_j),
//              f = g r^2 / (r^2 + eps);
//   XSPH       xs_i -= (m / rho_j) W_ij (v_i - v_j);
//   near       a_i += k_near (w_i + w_j) s(|r|) r, s = wn^2 / max(|r|, 1e-5),
//              wn = max(1 - |r| / h, 0), w the rho_near of density_div;
//   cohesion   a_i -= gamma corr_ij c(|r|) r, c = m C(|r|) / max(|r|, 1e-9),
//              corr_ij = 2 rho0 / (rho_i + rho_j);
// and the per-particle wall adhesion a_i += beta rho0 A(x_i) (a diagonal
// Jacobian). Writes lam_w (the adjoint of rho_near), adds to lam_v, lam_rho
// and lam_x, and d loss / d mu = 2 sum P_ij la_i . (v_i - v_j) (one atomic
// per block).
__global__ __launch_bounds__(kQuadBlock) void k_rev_forces(SceneK k, uint32_t n, SolveBufs b,
                                                           const float4* __restrict__ pos,
                                                           const float4* __restrict__ vel,
                                                           const float* __restrict__ rho,
                                                           const float* __restrict__ w_near,
                                                           const float4* __restrict__ lam_a,
                                                           const float4* __restrict__ lam_xs,
                                                           float4* __restrict__ lam_v, float* __restrict__ lam_rho,
                                                           float4* __restrict__ lam_x, float* __restrict__ lam_w,
                                                           float* __restrict__ lam_mu) {
    __shared__ float red[kQuadBlock / 64];
    const Quad q = quad(n, b);
    f3 s_v = mk(0.0f, 0.0f, 0.0f), s_x = s_v, la = s_v, xi = s_v;
    float s_r = 0.0f, s_w = 0.0f, s_mu = 0.0f, ri = 1.0f;
    if (q.live) {
        xi = ld3(pos, q.p);
        const f3 vi = ld3(vel, q.p), lxi = ld3(lam_xs, q.p);
        la = ld3(lam_a, q.p);
        ri = rho[q.p];
        const float wi = w_near[q.p];
        const float mri = div_rn(k.mass, ri);
        for_lane(b, q, [&](uint32_t j, float g) {
            const PairGeom P = pair_geom(k, xi, ld3(pos, j), g);
            const float r2 = P.r2, d = P.d;
            const float rj = rho[j], mrj = div_rn(k.mass, rj);
            const f3 dv = sub(vi, ld3(vel, j)), lxj = ld3(lam_xs, j);
            const f3 dl = sub(la, ld3(lam_a, j));
            // viscosity
            const float den = r2 + k.eps_h2;
            const float mrr = div_rn(k.mass, ri * rj);
            const float pair = mrr * div_rn(P.g * r2, den);
            const float dldv = dot3(dl, dv);
            s_v = add(s_v, scl(k.visc2 * pair, dl));
            s_r -= k.visc2 * pair * dldv / ri;
            const float fp_over_r = div_rn(P.hh * r2, den) + div_rn(2.0f * P.g * k.eps_h2, den * den);
            float sx = k.visc2 * mrr * fp_over_r * dldv;
            s_mu += 2.0f * pair * dot3(la, dv);
            // XSPH
            const float W = kernel_q(k, d * k.inv_h);
            s_v = add(s_v, scl(W, sub(scl(mri, lxj), scl(mrj, lxi))));
            const float lxjdv = dot3(lxj, dv);
            s_r -= mri / ri * W * lxjdv;
            sx += P.g * (mri * lxjdv - mrj * dot3(lxi, dv));
            // near pressure
            const float wn = wmax(1.0f - d * k.inv_h, 0.0f);
            if (wn > 0.0f) {
                const bool far = d >= 1e-5f;
                const float inv_rm = far ? 1.0f / d : 1.0f / 1e-5f;
                const float sn = wn * wn * inv_rm;
                const float snp = -2.0f * wn * k.inv_h * inv_rm - (far ? sn / d : 0.0f);  // ds / d|r|
                const f3 dln = scl(k.k_near, dl);
                const float xd = dot3(P.r, dln);
                s_w += sn * xd;
                const float wsum = wi + w_near[j];
                s_x = add(s_x, scl(wsum * sn, dln));
                if (d > 1e-9f) sx += wsum * snp / d * xd;
            }
            // cohesion
            const float C = cohesion(d, k.coh_c, k.coh_k, k.coh_near), Cp = cohesion_dr(k, d);
            if (C != 0.0f || Cp != 0.0f) {
                const float inv_rg = 1.0f / wmax(d, 1e-9f);
                const float c = k.mass * C * inv_rg;
                const float cp = k.mass * (Cp * inv_rg - C * inv_rg * inv_rg);  // dc / d|r|
                const float corr = div_rn(2.0f * k.rho0, ri + rj);
                const f3 dlt = scl(k.gamma, dl);
                const float xt = dot3(P.r, dlt);
                s_r += corr / (ri + rj) * c * xt;
                s_x = sub(s_x, scl(corr * c, dlt));
                sx -= corr * cp * inv_rg * xt;
            }
            s_x = add(s_x, scl(sx, P.r));
        });
    }
    s_v = fold4(s_v);
    s_x = fold4(s_x);
    s_r = fold4(s_r);
    s_w = fold4(s_w);
    s_mu = fold4(s_mu);
    float dmu = 0.0f;
    if (q.live && q.lane == 0) {
        const uint32_t i = q.p;
        // wall adhesion: A_a = -wa((x_a - lo_a) / h) + wa((hi_a - x_a) / h)
        if (near_walls(k, xi)) {
            const float ih = k.inv_h_wall, s_adh = k.beta * k.rho0;
            const float xa[3] = {xi.x, xi.y, xi.z}, lav[3] = {la.x, la.y, la.z};
            float ad[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float tl = (xa[a] - k.wall_lo[a]) * ih, th = (k.wall_hi[a] - xa[a]) * ih;
                ad[a] = -s_adh * ih * (wall_adhesion_dt(tl) + wall_adhesion_dt(th)) * lav[a];
            }
            s_x = add(s_x, mk(ad[0], ad[1], ad[2]));
        }
        add3(lam_v, i, s_v);
        lam_rho[i] += s_r;
        add3(lam_x, i, s_x);
        lam_w[i] = s_w;
        dmu = s_mu;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dmu += __shfl_xor(dmu, o);
    if (__lane_id() == 0) red[threadIdx.x / 64] = dmu;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (uint32_t w = 0; w < kQuadBlock / 64; ++w) t += red[w];
        atomicAdd(lam_mu, t);
    }
}

// density_div in reverse, part 1 (per particle, with a sweep recomputing G =
// m sum gradW + wg and gsq = m^2 sum |gradW|^2): kd = max(D, 0) alpha / dt /
// rho and alpha = rho / max(|G|^2 + gsq, 1e-4) give lam_D, lam_G, lam_gsq and
// the final lam_rho.
__global__ __launch_bounds__(kQuadBlock) void k_rev_density_pre(SceneK k, StepK s, uint32_t n, SolveBufs b,
                                                                const float4* __restrict__ pos,
                                                                const float4* __restrict__ wg,
                                                                const float* __restrict__ rho,
                                                                const float* __restrict__ alpha,
                                                                const float* __restrict__ kd,
                                                                const float* __restrict__ lam_kd,
                                                                float* __restrict__ lam_alpha,
                                                                float* __restrict__ lam_rho,
                                                                float* __restrict__ lam_D,
                                                                float4* __restrict__ lam_G) {
    const Quad q = quad(n, b);
    f3 G = mk(0.0f, 0.0f, 0.0f);
    float gsq = 0.0f;
    if (q.live) {
        const f3 xi = ld3(pos, q.p);
        for_lane(b, q, [&](uint32_t j, float g) {
            const f3 gw = scl(g, sub(xi, ld3(pos, j)));
            G = add(G, gw);
            gsq += dot3(gw, gw);
        });
    }
    G = fold4(G);
    gsq = fold4(gsq);
    if (!q.live || q.lane != 0) return;
    const uint32_t i = q.p;
    G = add(scl(k.mass, G), ld3(wg, i));
    gsq *= k.mass * k.mass;
    const float r = rho[i], a = alpha[i], kk = kd[i], lk = lam_kd[i];
    float la = lam_alpha[i], lr = lam_rho[i];
    float lD = 0.0f;
    if (kk > 0.0f) {
        lD = lk * a / (s.dt_guard * r);
        la += lk * kk / a;
    }
    lr += -lk * kk / r;
    const float qq = dot3(G, G) + gsq;
    const float den = wmax(qq, 1e-4f);
    lr += la / den;
    float lden = 0.0f;
    if (qq > 1e-4f) lden = -la * a / den;
    lam_rho[i] = lr;
    lam_D[i] = lD;
    lam_G[i] = make_float4(2.0f * lden * G.x, 2.0f * lden * G.y, 2.0f * lden * G.z, lden);  // w: lam_gsq
}

// density_div in reverse, part 2 (the sweep): lam_v from D, lam_x from D, G,
// gsq, rho, rho_near and the walls. rho = m (sigma + sum W) + rho0 fill(x);
// rho_near w_i = sum wn_ij^3, wn = max(1 - |r| / h, 0), whose adjoint lam_w
// forces_eval's reverse wrote.
__global__ __launch_bounds__(kQuadBlock) void k_rev_density(SceneK k, uint32_t n, SolveBufs b,
                                                            const float4* __restrict__ pos,
                                                            const float4* __restrict__ vel,
                                                            const float4* __restrict__ wg,
                                                            const float4* __restrict__ J,
                                                            const float4* __restrict__ dfill,
                                                            const float* __restrict__ lam_rho,
                                                            const float* __restrict__ lam_D,
                                                            const float4* __restrict__ lam_G,
                                                            const float* __restrict__ lam_w,
                                                            float4* __restrict__ lam_v, float4* __restrict__ lam_x) {
    const Quad q = quad(n, b);
    f3 s_v = mk(0.0f, 0.0f, 0.0f), s_x = s_v, s_n = s_v, vi = s_v, Gi = s_v;
    float di = 0.0f, ri = 0.0f;
    if (q.live) {
        const f3 xi = ld3(pos, q.p);
        vi = ld3(vel, q.p);
        di = lam_D[q.p];
        ri = lam_rho[q.p];
        const float wi = lam_w[q.p];
        const float4 Gi4 = lam_G[q.p];
        Gi = xyz(Gi4);
        const float si = Gi4.w;
        for_lane(b, q, [&](uint32_t j, float g) {
            const PairGeom P = pair_geom(k, xi, ld3(pos, j), g);
            const f3 gw = scl(P.g, P.r);
            const float dj = lam_D[j];
            const float4 Gj4 = lam_G[j];
            s_v = add(s_v, scl(di + dj, gw));
            f3 t = scl(di + dj, hess(P, sub(vi, ld3(vel, j))));
            t = add(t, hess(P, sub(Gi, xyz(Gj4))));
            t = add(t, scl(2.0f * k.mass * (si + Gj4.w), hess(P, gw)));
            t = add(t, scl(ri + lam_rho[j], gw));
            s_x = add(s_x, t);
            const float wn = wmax(1.0f - P.d * k.inv_h, 0.0f);
            if (wn > 0.0f && P.d > 1e-9f)
                s_n = add(s_n, scl((wi + lam_w[j]) * (-3.0f * wn * wn * k.inv_h / P.d), P.r));
        });
    }
    s_v = fold4(s_v);
    s_x = fold4(s_x);
    s_n = fold4(s_n);
    if (!q.live || q.lane != 0) return;
    const uint32_t i = q.p;
    add3(lam_v, i, add(scl(k.mass, s_v), scl(di, ld3(wg, i))));
    f3 w = add(scl(di, jt(J, i, vi)), jt(J, i, Gi));
    w = add(w, add(scl(k.rho0 * ri, ld3(dfill, i)), s_n));
    add3(lam_x, i, add(scl(k.mass, s_x), w));
}

// The working set came from the resting set through the permutation: the
// adjoints go back to the previous substep's resting order.
__global__ __launch_bounds__(kAdjBlock) void k_rev_gather(uint32_t n, const uint32_t* __restrict__ perm,
                                                          const float4* __restrict__ lx, const float4* __restrict__ lv,
                                                          const float* __restrict__ lpp, float4* __restrict__ ox,
                                                          float4* __restrict__ ov, float* __restrict__ opp) {
    const uint32_t f = blockIdx.x * kAdjBlock + threadIdx.x;
    if (f >= n) return;
    const uint32_t i = perm[f];
    ox[i] = lx[f];
    ov[i] = lv[f];
    opp[i] = lpp[f];
}

__global__ __launch_bounds__(kAdjBlock) void k_fill_zero(float* __restrict__ a, size_t count) {
    const size_t i = (size_t)blockIdx.x * kAdjBlock + threadIdx.x;
    if (i < count) a[i] = 0.0f;
}

}  // namespace

static dim3 adj_grid(uint32_t n) { return dim3((n + kAdjBlock - 1) / kAdjBlock); }
static dim3 quad_grid(uint32_t n) { return dim3((n * kListLanes + kQuadBlock - 1) / kQuadBlock); }

void adj_zero(hipStream_t st, float* a, size_t count) {
    if (count) hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((count + kAdjBlock - 1) / kAdjBlock)),
                                  dim3(kAdjBlock), 0, st, a, count);
}

// One substep in reverse. On entry L.x, L.v hold the adjoints of the resting
// position and velocity integrate wrote (working slots), L.p that of the
// pressure it carried over; on exit L.x, L.v, L.pp hold the adjoints of the
// previous resting set (its order), and L.g / L.mu have this substep's
// gravity and viscosity terms added.
void launch_reverse_substep(hipStream_t st, const SceneK& k, const StepK& s, uint32_t n, uint32_t passes,
                            const Tape& T, const SolveBufs& b, const AdjointBufs& L) {
    if (n == 0) return;
    const dim3 g = adj_grid(n), blk(kAdjBlock), gq = quad_grid(n), bq(kQuadBlock);
    hipLaunchKernelGGL(k_rev_integrate, g, blk, 0, st, k, s, n, T.pos, T.vel[passes + 2], T.density, L.x, L.v,
                       L.dfill, L.J, L.rho, L.alpha);
    for (uint32_t r = passes; r-- > 0;) {
        // den_apply r (kd of pass r, lam of its output = L.v), then den_kappa
        // r's per-particle reverse and its sweep at the velocities entering
        // pass r
        hipLaunchKernelGGL(k_rev_apply<true>, gq, bq, 0, st, k, s, n, b, T.pos, T.kd[2 + r], T.wall_grad, L.J, L.v,
                           L.kd, L.x, T.density, T.alpha, L.p, L.D, L.rho, L.alpha);
        hipLaunchKernelGGL(k_rev_divergence, gq, bq, 0, st, k, n, b, T.pos, T.vel[2 + r], T.wall_grad, L.J, L.D,
                           L.v, L.x);
    }
    // forces_den_apply: the warm-start correction, then v2 = v1 + dt (g + a)
    hipLaunchKernelGGL(k_rev_apply<false>, gq, bq, 0, st, k, s, n, b, T.pos, T.kd[1], T.wall_grad, L.J, L.v, L.kd,
                       L.x, T.density, T.alpha, L.p, L.D, L.rho, L.alpha);
    hipLaunchKernelGGL(k_rev_forces_den, g, blk, 0, st, s, n, T.density, T.pp, L.v, L.kd, L.p, L.a, L.xs, L.x,
                       L.pp, L.rho, L.g);
    // forces_eval at the velocities after div_apply
    hipLaunchKernelGGL(k_rev_forces, gq, bq, 0, st, k, n, b, T.pos, T.vel[1], T.density, T.rho_near, L.a, L.xs, L.v,
                       L.rho, L.x, L.w, L.mu);
    // div_apply with the divergence kappa
    hipLaunchKernelGGL(k_rev_apply<false>, gq, bq, 0, st, k, s, n, b, T.pos, T.kd[0], T.wall_grad, L.J, L.v, L.kd,
                       L.x, T.density, T.alpha, L.p, L.D, L.rho, L.alpha);
    // density_div at the substep's first velocities
    hipLaunchKernelGGL(k_rev_density_pre, gq, bq, 0, st, k, s, n, b, T.pos, T.wall_grad, T.density, T.alpha,
                       T.kd[0], L.kd, L.alpha, L.rho, L.D, L.G);
    hipLaunchKernelGGL(k_rev_density, gq, bq, 0, st, k, n, b, T.pos, T.vel[0], T.wall_grad, L.J, L.dfill, L.rho,
                       L.D, L.G, L.w, L.v, L.x);
    hipLaunchKernelGGL(k_rev_gather, g, blk, 0, st, n, T.perm, L.x, L.v, L.pp, L.x_out, L.v_out, L.pp_out);
}

}  // namespace fluid
