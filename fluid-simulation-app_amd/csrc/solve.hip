// The DFSPH solve for gfx950: density and the DFSPH factor, the divergence
// and constant-density pressure solves, the force terms, and integrate.
//
// Replaces the reference's sweeps (crates/fluid-core/src/sim_solve.wgsl):
//   density_div        411-518   27-cell walk; density, near-density, alpha,
//                                divergence kappa, wall gradient, neighbour list
//   div_apply          665-668   (kappa_partial + apply_kappa_wide, 594-621)
//   forces_eval        524-592
//   forces_den_apply   630-663
//   den_kappa          674-698
//   den_apply          700-703
//   integrate          709-756
//
// Layout: every array is in slot (cell) order after the grid build; one
// thread owns one particle and walks its neighbours serially in the
// reference's cell numbering (x fastest, then y, then z; in-cell slot order),
// which is the reference's LANES = 1 reading. Each x-row of the 3x3x3
// stencil is one contiguous slot range, so the walk reads 9 ranges.
//
// density_div walks the stencil in two phases: a distance test per candidate
// that appends accepted slots to the neighbour list, then the per-pair math
// over that compact list (so the wave does not run the heavy path on every
// candidate some lane accepts). The list holds, per pair, the slot and the
// scalar dW/r (neighbour-major, so a wave's reads of entry m are one
// coalesced line); the later sweeps rebuild each gradient as dW/r times
// x_i - x_j, the product grad_q itself forms, so it is bit-identical, with
// 8 bytes a pair from HBM instead of the reference's 16 and no square root
// or division.
#include "fluid_common.hpp"
#include "fluid_kernels.hpp"

namespace fluid {

namespace {

__device__ __forceinline__ uint32_t particle_index() {
    return xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
}

// The gradient of a listed pair: the stored dW/r times x_i - x_j, which is
// grad_q's own product (sim_solve.wgsl:195), so it is bit-identical to the
// gradient density_div computed.
__device__ __forceinline__ f3 listed_grad(f3 pi, float4 pj, float g) {
    return scl(g, sub(pi, xyz(pj)));
}

// The heavy per-pair terms of density_div for one accepted neighbour.
struct DensitySums {
    float w_sum, rn, gsq, drho;
    f3 gsum;
};

__device__ __forceinline__ float density_pair(const SceneK& k, f3 pi, f3 vi, float4 pj, float4 vj, DensitySums& a) {
    const f3 x = sub(pi, xyz(pj));
    const float r2 = dot3(x, x);
    const float r = sqrtf(r2);
    const float q = r * k.inv_h;
    a.w_sum += kernel_q(k, q);
    const float wn = wmax(1.0f - q, 0.0f);
    a.rn += wn * wn * wn;
    const float g = r < 1e-9f ? 0.0f : grad_scale(k, r, q);
    const f3 gw = r < 1e-9f ? mk(0.0f, 0.0f, 0.0f) : scl(g, x);
    a.gsum = add(a.gsum, gw);
    a.gsq += dot3(gw, gw);
    a.drho += dot3(sub(vi, xyz(vj)), gw);
    return g;
}

// The candidate rows of the 3x3x3 stencil around cell (bx, by, bz): row
// (dy, dz) spans x cells [x0, x1]. A row or an end cell whose box lies
// farther from the particle than the support is left out: no particle in it
// can pass the r^2 < support^2 test, so the accepted neighbours, and their
// order, are those of the full 27-cell walk (sim_solve.wgsl:436-468).
struct Stencil {
    float dl[3], dh[3];  // distance to the low / high face of the own cell, per axis
    float reach;         // support^2 with a margin over the rounding of the faces
};

__device__ __forceinline__ Stencil make_stencil(const SceneK& k, f3 p, uint32_t bx, uint32_t by, uint32_t bz) {
    Stencil st;
    const float pp[3] = {p.x, p.y, p.z};
    const uint32_t b[3] = {bx, by, bz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float lo = k.box_min[a] + (float)b[a] * k.cell;
        const float hi = k.box_min[a] + (float)(b[a] + 1u) * k.cell;
        st.dl[a] = wmax(pp[a] - lo, 0.0f);
        st.dh[a] = wmax(hi - pp[a], 0.0f);
    }
    st.reach = k.support_sq * 1.001f;
    return st;
}

__global__ __launch_bounds__(kBlock) void k_density_div(SceneK k, StepK s, uint32_t n, const uint32_t* __restrict__ starts,
                                                        const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                        SolveBufs b) {
    uint32_t i = particle_index();
    if (i >= n) return;
    const f3 pi = xyz(pos[i]);
    const f3 vi = xyz(vel[i]);
    const uint32_t bx = cell_coord(pi.x, k.box_min[0], k.cell, k.dims[0]);
    const uint32_t by = cell_coord(pi.y, k.box_min[1], k.cell, k.dims[1]);
    const uint32_t bz = cell_coord(pi.z, k.box_min[2], k.cell, k.dims[2]);
    const Stencil st = make_stencil(k, pi, bx, by, bz);
    uint2* __restrict__ list = b.nbr + i;
    const uint32_t stride = b.stride, cap = b.cap;
    // Phase 1: the candidate walk, in the reference's cell order, appending
    // accepted slots to the list (a distance test per candidate and nothing
    // else, so the wave stays converged).
    uint32_t found = 0;
    for (int dz = -1; dz <= 1; ++dz) {
        const int cz = (int)bz + dz;
        if (cz < 0 || cz >= (int)k.dims[2]) continue;
        const float ez = dz < 0 ? st.dl[2] : (dz > 0 ? st.dh[2] : 0.0f);
        for (int dy = -1; dy <= 1; ++dy) {
            const int cy = (int)by + dy;
            if (cy < 0 || cy >= (int)k.dims[1]) continue;
            const float ey = dy < 0 ? st.dl[1] : (dy > 0 ? st.dh[1] : 0.0f);
            const float e2 = ey * ey + ez * ez;
            if (e2 > st.reach) continue;
            const uint32_t x0 = (bx > 0 && st.dl[0] * st.dl[0] + e2 <= st.reach) ? bx - 1 : bx;
            const uint32_t x1 = (bx + 1 < k.dims[0] && st.dh[0] * st.dh[0] + e2 <= st.reach) ? bx + 1 : bx;
            const uint32_t row = ((uint32_t)cz * k.dims[1] + (uint32_t)cy) * k.dims[0];
            const uint32_t k0 = starts[row + x0], k1 = starts[row + x1 + 1];
            for (uint32_t kk = k0; kk < k1; ++kk) {
                const f3 x = sub(pi, xyz(pos[kk]));
                const float r2 = dot3(x, x);
                if (r2 < k.support_sq && kk != i) {
                    if (found < cap) list[(size_t)found * stride].x = kk;
                    found++;
                }
            }
        }
    }
    // Phase 2: the sums over the compact list, in list order.
    DensitySums a{k.sigma, 0.0f, 0.0f, 0.0f, mk(0.0f, 0.0f, 0.0f)};
    const uint32_t nn = found < cap ? found : cap;
    for (uint32_t m = 0; m < nn; ++m) {
        const uint32_t j = list[(size_t)m * stride].x;
        const float g = density_pair(k, pi, vi, pos[j], vel[j], a);
        list[(size_t)m * stride].y = __float_as_uint(g);
    }
    if (found > cap) {
        // Past the cap the list keeps its first `cap` neighbours (the later
        // sweeps see only those, as in the reference) but the density sums
        // take every neighbour: walk again for the ones the list dropped.
        atomicAdd(&b.counters[1], found - cap);
        uint32_t seen = 0;
        for (int dz = -1; dz <= 1; ++dz) {
            const int cz = (int)bz + dz;
            if (cz < 0 || cz >= (int)k.dims[2]) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                const int cy = (int)by + dy;
                if (cy < 0 || cy >= (int)k.dims[1]) continue;
                const uint32_t x0 = bx > 0 ? bx - 1 : bx;
                const uint32_t x1 = bx + 1 < k.dims[0] ? bx + 1 : bx;
                const uint32_t row = ((uint32_t)cz * k.dims[1] + (uint32_t)cy) * k.dims[0];
                const uint32_t k0 = starts[row + x0], k1 = starts[row + x1 + 1];
                for (uint32_t kk = k0; kk < k1; ++kk) {
                    const f3 x = sub(pi, xyz(pos[kk]));
                    if (dot3(x, x) < k.support_sq && kk != i) {
                        if (seen >= cap) density_pair(k, pi, vi, pos[kk], vel[kk], a);
                        seen++;
                    }
                }
            }
        }
    }
    b.nbr_n[i] = nn;
    float rho = k.mass * a.w_sum;
    const float gsq = k.mass * k.mass * a.gsq;
    float drho = k.mass * a.drho;
    f3 gsum = scl(k.mass, a.gsum);
    const f3 tl = mk((pi.x - k.wall_lo[0]) * k.inv_h, (pi.y - k.wall_lo[1]) * k.inv_h, (pi.z - k.wall_lo[2]) * k.inv_h);
    const f3 th = mk((k.wall_hi[0] - pi.x) * k.inv_h, (k.wall_hi[1] - pi.y) * k.inv_h, (k.wall_hi[2] - pi.z) * k.inv_h);
    float fill = 0.0f;
    fill += wall_density(tl.x) + wall_density(th.x);
    fill += wall_density(tl.y) + wall_density(th.y);
    fill += wall_density(tl.z) + wall_density(th.z);
    fill -= wedge_sum(tl, th);
    rho += k.rho0 * fill;
    const f3 wg = wall_grad_sum(k, pi);
    b.wall_grad[i] = make_float4(wg.x, wg.y, wg.z, 0.0f);
    gsum = add(gsum, wg);
    b.density[i] = rho;
    b.rho_near[i] = a.rn;
    const float al = rho / wmax(dot3(gsum, gsum) + gsq, 1e-4f);
    b.alpha[i] = al;
    drho += dot3(vi, wg);
    b.kd[i] = wmax(drho, 0.0f) * al / s.dt_guard / rho;
}

// div_apply / den_apply: v -= dt (m sum (kd_i + kd_j) gradW + kd_i wall_grad).
// Reads only kd of the neighbours, so the velocity update is in place.
__global__ __launch_bounds__(kBlock) void k_kappa_apply(SceneK k, StepK s, uint32_t n, const float4* __restrict__ pos,
                                                        float4* __restrict__ vel, SolveBufs b) {
    uint32_t i = particle_index();
    if (i >= n) return;
    const f3 pi = xyz(pos[i]);
    const float kdi = b.kd[i];
    const uint32_t nn = b.nbr_n[i];
    f3 dv = mk(0.0f, 0.0f, 0.0f);
    const uint2* __restrict__ list = b.nbr + i;
#pragma unroll 4
    for (uint32_t m = 0; m < nn; ++m) {
        const uint2 e = list[(size_t)m * b.stride];
        dv = add(dv, scl(kdi + b.kd[e.x], listed_grad(pi, pos[e.x], __uint_as_float(e.y))));
    }
    const float4 wg = b.wall_grad[i];
    const f3 full = add(scl(k.mass, dv), scl(kdi, xyz(wg)));
    const float4 w = vel[i];
    vel[i] = make_float4(w.x - s.dt * full.x, w.y - s.dt * full.y, w.z - s.dt * full.z, w.w);
}

__global__ __launch_bounds__(kBlock) void k_forces_eval(SceneK k, uint32_t n, const float4* __restrict__ pos,
                                                        const float4* __restrict__ vel,
                                                        const float* __restrict__ temperature,
                                                        const float* __restrict__ prev_pressure, SolveBufs b) {
    uint32_t i = particle_index();
    if (i >= n) return;
    const f3 pi = xyz(pos[i]);
    const f3 vi = xyz(vel[i]);
    const float w_i = b.rho_near[i], temp_i = temperature[i], rho_i = b.density[i];
    f3 visc = mk(0.0f, 0.0f, 0.0f), blend = visc, nearf = visc, tension = visc;
    float heat = 0.0f;
    const uint32_t nn = b.nbr_n[i];
    const uint2* __restrict__ list = b.nbr + i;
    for (uint32_t m = 0; m < nn; ++m) {
        const uint2 e = list[(size_t)m * b.stride];
        const uint32_t j = e.x;
        const f3 x = sub(pi, xyz(pos[j]));
        const float r2 = dot3(x, x);
        const float r = sqrtf(r2);
        const f3 g = scl(__uint_as_float(e.y), x);
        const float rho_j = b.density[j];
        const float f = dot3(x, g) / (r2 + k.eps_h2);
        const float pair = k.mass / (rho_i * rho_j) * f;
        const f3 dv = sub(vi, xyz(vel[j]));
        const float wn = wmax(1.0f - r * k.inv_h, 0.0f);
        const float rm = wmax(r, 1e-5f);
        const float sn = (w_i + b.rho_near[j]) * wn * wn;
        nearf = add(nearf, mk(sn * (x.x / rm), sn * (x.y / rm), sn * (x.z / rm)));
        visc = add(visc, scl(k.visc2 * pair, dv));
        blend = sub(blend, scl(k.mass / rho_j * kernel_q(k, r * k.inv_h), dv));
        heat += k.heat_cond * pair * (temp_i - temperature[j]);
        heat -= k.heat_visc * pair * dot3(dv, dv);
        const float corr = 2.0f * k.rho0 / (rho_i + rho_j);
        const float cm = corr * k.mass * cohesion(r, k.coh_c, k.coh_k, k.coh_near);
        const float rg = wmax(r, 1e-9f);
        tension = sub(tension, mk(cm * x.x / rg, cm * x.y / rg, cm * x.z / rg));
    }
    const f3 adh = scl(k.beta * k.rho0, wall_adh_sum(k, pi));
    const f3 acc = add(add(add(visc, scl(K_NEAR, nearf)), scl(k.gamma, tension)), adh);
    b.accel[i] = make_float4(acc.x, acc.y, acc.z, heat);
    b.xsph[i] = make_float4(blend.x, blend.y, blend.z, 0.0f);
    const float half = 0.5f * prev_pressure[i];
    b.kd[i] = half / (rho_i * rho_i);
    b.pressure[i] = half;
}

// sim_solve.wgsl:103-127
__device__ __forceinline__ f3 fingers(const StepK& s, f3 pos, f3 v) {
    float pullx = 0.0f, pully = 0.0f, weight = 0.0f;
    for (uint32_t t = 0; t < s.touch_n; ++t) {
        const float dx = pos.x - s.touches[t][0], dy = pos.y - s.touches[t][1];
        const float q2 = (dx * dx + dy * dy) / (s.touch_r * s.touch_r);
        if (q2 < 1.0f) {
            const float w = (1.0f - q2) * (1.0f - q2);
            pullx += w * s.touches[t][2];
            pully += w * s.touches[t][3];
            weight += w;
        }
    }
    if (weight <= 0.0f) return v;
    const float blend = 1.0f - exp_fixed(-TOUCH_RATE * weight * s.dt);
    return mk(wmix(v.x, pullx / weight, blend), wmix(v.y, pully / weight, blend), v.z);
}

__global__ __launch_bounds__(kBlock) void k_forces_den_apply(SceneK k, StepK s, uint32_t n,
                                                             const float4* __restrict__ pos, float4* __restrict__ vel,
                                                             float* __restrict__ temperature, SolveBufs b) {
    uint32_t i = particle_index();
    if (i >= n) return;
    const f3 pi = xyz(pos[i]);
    const float kdi = b.kd[i];
    const uint32_t nn = b.nbr_n[i];
    f3 dv = mk(0.0f, 0.0f, 0.0f);
    const uint2* __restrict__ list = b.nbr + i;
#pragma unroll 4
    for (uint32_t m = 0; m < nn; ++m) {
        const uint2 e = list[(size_t)m * b.stride];
        dv = add(dv, scl(kdi + b.kd[e.x], listed_grad(pi, pos[e.x], __uint_as_float(e.y))));
    }
    const float4 a = b.accel[i];
    const float4 was = vel[i];
    const f3 v = xyz(was);
    const f3 om = mk(s.omega[0], s.omega[1], s.omega[2]);
    const f3 dom = mk(s.domega[0], s.domega[1], s.domega[2]);
    const f3 c1 = cross3(dom, pi);
    const f3 c2 = cross3(om, cross3(om, pi));
    const f3 c3 = scl(2.0f, cross3(om, v));
    const f3 fict = sub(sub(mk(-c1.x, -c1.y, -c1.z), c2), c3);
    const f3 fsum = add(add(mk(s.force[0], s.force[1], s.force[2]), fict), xyz(a));
    const float4 xs = b.xsph[i];
    const f3 forced = fingers(s, pi, add(add(v, scl(s.dt, fsum)), scl(s.c_xsph, xyz(xs))));
    temperature[i] = temperature[i] + s.dt * a.w;
    const float4 wg = b.wall_grad[i];
    const f3 full = add(scl(k.mass, dv), scl(kdi, xyz(wg)));
    vel[i] = make_float4(forced.x - s.dt * full.x, forced.y - s.dt * full.y, forced.z - s.dt * full.z, was.w);
}

__global__ __launch_bounds__(kBlock) void k_den_kappa(SceneK k, StepK s, uint32_t n, const float4* __restrict__ pos,
                                                      const float4* __restrict__ vel, SolveBufs b) {
    uint32_t i = particle_index();
    if (i >= n) return;
    const f3 pi = xyz(pos[i]);
    const f3 vi = xyz(vel[i]);
    const uint32_t nn = b.nbr_n[i];
    float drho = 0.0f;
    const uint2* __restrict__ list = b.nbr + i;
#pragma unroll 4
    for (uint32_t m = 0; m < nn; ++m) {
        const uint2 e = list[(size_t)m * b.stride];
        drho += dot3(sub(vi, xyz(vel[e.x])), listed_grad(pi, pos[e.x], __uint_as_float(e.y)));
    }
    const float rho = b.density[i];
    const float4 wg = b.wall_grad[i];
    drho = k.mass * drho + dot3(vi, xyz(wg));
    const float rho_star = rho + s.dt * drho;
    const float kap = wmax(rho_star - k.rho0, 0.0f) * b.alpha[i] / (s.dt * s.dt);
    b.kd[i] = kap / rho;
    b.pressure[i] = b.pressure[i] + kap * rho;
}

// Reads the working set, writes the resting set at the same slot.
__global__ __launch_bounds__(kBlock) void k_integrate(SceneK k, StepK s, uint32_t n, Records work, Records rest,
                                                      SolveBufs b) {
    uint32_t i = particle_index();
    bool clamped_speed = false;
    if (i < n) {
        const float4 was = work.vel[i];
        f3 v = xyz(was);
        const float speed = sqrtf(dot3(v, v));
        const float rho = b.density[i];
        const float ceiling = rho < 0.5f * k.rho0 ? 3.0f * s.v_clamp : s.v_clamp;
        if (speed > ceiling) {
            v = scl(ceiling / speed, v);
            clamped_speed = true;
        }
        const float4 pv = work.prev_vel[i];
        const f3 dvv = sub(v, xyz(pv));
        const float acc = sqrtf(dot3(dvv, dvv)) / s.dt;
        const float4 p4 = work.pos[i];
        const float in[3] = {p4.x + s.dt * v.x, p4.y + s.dt * v.y, p4.z + s.dt * v.z};
        float vv[3] = {v.x, v.y, v.z};
        float out[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float c = wmin(wmax(in[a], k.wall_lo[a]), k.wall_hi[a]);
            out[a] = c;
            if (c != in[a]) vv[a] = 0.0f;
        }
        const float pr = b.pressure[i];
        rest.pos[i] = make_float4(out[0], out[1], out[2], 0.0f);
        rest.vel[i] = make_float4(vv[0], vv[1], vv[2], wmix(was.w, acc, s.chase));
        rest.prev_vel[i] = make_float4(vv[0], vv[1], vv[2], wmix(pv.w, pr, s.chase));
        const float dp = pr - work.prev_pressure[i];
        rest.prev_pressure[i] = pr;
        const float t = work.temperature[i];
        rest.temperature[i] = t + t * EXPANSION / (rho * HEAT_CAPACITY) * dp;
        if (rest.ids) rest.ids[i] = work.ids[i];
    }
    // One atomic per wave for the clamp counter.
    const uint64_t m = __ballot(clamped_speed);
    if (m && __lane_id() == (uint32_t)(__ffsll((unsigned long long)m) - 1))
        atomicAdd(&b.counters[0], (uint32_t)__popcll(m));
}

}  // namespace

void launch_density_only(hipStream_t st, const SceneK& k, const StepK& s, uint32_t n, const uint32_t* starts,
                         const Records& work, const SolveBufs& b, KernelTimer* timer) {
    if (n == 0) return;
    const dim3 grid(blocks_for(n)), block(kBlock);
    if (timer) timer->begin(st, KERNEL_DENSITY_DIV);
    hipLaunchKernelGGL(k_density_div, grid, block, 0, st, k, s, n, starts, work.pos, work.vel, b);
    if (timer) timer->end(st, KERNEL_DENSITY_DIV);
}

void launch_solve(hipStream_t st, const SceneK& k, const StepK& s, uint32_t n, const uint32_t* starts,
                  const Records& work, const Records& rest, const SolveBufs& b, uint32_t refine,
                  KernelTimer* timer) {
    if (n == 0) return;
    const dim3 grid(blocks_for(n)), block(kBlock);
    launch_density_only(st, k, s, n, starts, work, b, timer);
    if (timer) timer->begin(st, KERNEL_DIV_APPLY);
    hipLaunchKernelGGL(k_kappa_apply, grid, block, 0, st, k, s, n, work.pos, work.vel, b);
    if (timer) timer->end(st, KERNEL_DIV_APPLY);
    if (timer) timer->begin(st, KERNEL_FORCES_EVAL);
    hipLaunchKernelGGL(k_forces_eval, grid, block, 0, st, k, n, work.pos, work.vel, work.temperature,
                       work.prev_pressure, b);
    if (timer) timer->end(st, KERNEL_FORCES_EVAL);
    if (timer) timer->begin(st, KERNEL_FORCES_DEN_APPLY);
    hipLaunchKernelGGL(k_forces_den_apply, grid, block, 0, st, k, s, n, work.pos, work.vel, work.temperature, b);
    if (timer) timer->end(st, KERNEL_FORCES_DEN_APPLY);
    for (uint32_t r = 0; r < refine; ++r) {
        if (timer) timer->begin(st, KERNEL_DEN_KAPPA);
        hipLaunchKernelGGL(k_den_kappa, grid, block, 0, st, k, s, n, work.pos, work.vel, b);
        if (timer) timer->end(st, KERNEL_DEN_KAPPA);
        if (timer) timer->begin(st, KERNEL_DEN_APPLY);
        hipLaunchKernelGGL(k_kappa_apply, grid, block, 0, st, k, s, n, work.pos, work.vel, b);
        if (timer) timer->end(st, KERNEL_DEN_APPLY);
    }
    if (timer) timer->begin(st, KERNEL_INTEGRATE);
    hipLaunchKernelGGL(k_integrate, grid, block, 0, st, k, s, n, work, rest, b);
    if (timer) timer->end(st, KERNEL_INTEGRATE);
}

}  // namespace fluid
