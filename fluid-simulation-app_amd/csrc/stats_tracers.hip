// Frame statistics and the passive tracers for gfx950.
//
// reduce_stats replaces the reference's single-workgroup reduction
// (crates/fluid-core/src/sim_solve.wgsl:800-874) by a device-wide one: a
// grid of workgroups folds strided slices into per-workgroup partials, and
// one workgroup folds the partials in workgroup order. Min and max are exact;
// the compression sum is a float sum in a different association than the
// reference's 256-way strided sum.
//
// The tracers follow sim_tracers.wgsl: splat (98-114) sums particle
// velocities per cell in 16.16 fixed point with integer atomics (exact in
// any order), resolve (122-138) copies and clears that grid, advect (140-183)
// samples it trilinearly and moves, recycles and repacks each tracer.
#include "fluid_common.hpp"
#include "fluid_kernels.hpp"

namespace fluid {

namespace {

constexpr int kStatsMaxBlocks = 1024;

// Workgroups of the stats fold for n particles (at most kStatsMaxBlocks).
__host__ __device__ __forceinline__ uint32_t stats_block_count(uint32_t n) {
    uint32_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    return b < (uint32_t)kStatsMaxBlocks ? b : (uint32_t)kStatsMaxBlocks;
}

__device__ __forceinline__ float lane_min(float a, float b) { return wmin(a, b); }
__device__ __forceinline__ float lane_max(float a, float b) { return wmax(a, b); }

// Workgroup b of `blocks` folds slots base + b*256 + t, base + (b + blocks)*256
// + t, ... below base + n into partials[b]. The batched form runs one row of
// workgroups per scene (blockIdx.y), each over its scene's own slots with the
// block count the plain step would take for that particle count.
__device__ __forceinline__ void stats_partial(float rho0, uint32_t base, uint32_t n, uint32_t blocks, uint32_t b,
                                              const float* __restrict__ density, const float4* __restrict__ vel,
                                              const float4* __restrict__ prev_vel,
                                              const float* __restrict__ temperature, float* __restrict__ out);

__global__ __launch_bounds__(kBlock) void k_stats_partial(float rho0, uint32_t n, const float* __restrict__ density,
                                                          const float4* __restrict__ vel,
                                                          const float4* __restrict__ prev_vel,
                                                          const float* __restrict__ temperature,
                                                          float* __restrict__ partials) {
    stats_partial(rho0, 0, n, gridDim.x, blockIdx.x, density, vel, prev_vel, temperature,
                  partials + (size_t)blockIdx.x * 16);
}

__global__ __launch_bounds__(kBlock) void k_stats_partial_many(SceneTable sc, const float* __restrict__ density,
                                                               const float4* __restrict__ vel,
                                                               const float4* __restrict__ prev_vel,
                                                               const float* __restrict__ temperature,
                                                               float* __restrict__ partials) {
    const SceneK& k = sc.ks[blockIdx.y];
    const uint32_t n = k.slot_end - k.slot_begin;
    const uint32_t blocks = stats_block_count(n);
    if (blockIdx.x >= blocks) return;
    stats_partial(k.rho0, k.slot_begin, n, blocks, blockIdx.x, density, vel, prev_vel, temperature,
                  partials + ((size_t)blockIdx.y * kStatsMaxBlocks + blockIdx.x) * 16);
}

__device__ __forceinline__ void stats_partial(float rho0, uint32_t base, uint32_t n, uint32_t blocks, uint32_t b,
                                              const float* __restrict__ density, const float4* __restrict__ vel,
                                              const float4* __restrict__ prev_vel,
                                              const float* __restrict__ temperature, float* __restrict__ out) {
    float acc[12] = {0.0f, 0.0f, HUGE_F, -HUGE_F, HUGE_F, -HUGE_F, HUGE_F, -HUGE_F, HUGE_F, -HUGE_F, HUGE_F, -HUGE_F};
    const uint32_t stride = blocks * kBlock;
    for (uint32_t li = b * kBlock + threadIdx.x; li < n; li += stride) {
        const uint32_t i = base + li;
        const float d = density[i];
        const float c = d / rho0 - 1.0f;
        acc[0] += wmax(c, 0.0f);
        acc[1] = wmax(acc[1], c);
        acc[2] = wmin(acc[2], d);
        acc[3] = wmax(acc[3], d);
        const float4 q = vel[i];
        const float sp = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z);
        acc[4] = wmin(acc[4], sp);
        acc[5] = wmax(acc[5], sp);
        acc[6] = wmin(acc[6], q.w);
        acc[7] = wmax(acc[7], q.w);
        const float pw = prev_vel[i].w;
        acc[8] = wmin(acc[8], pw);
        acc[9] = wmax(acc[9], pw);
        const float t = temperature[i];
        acc[10] = wmin(acc[10], t);
        acc[11] = wmax(acc[11], t);
    }
    __shared__ float red[12][kBlock];
#pragma unroll
    for (int f = 0; f < 12; ++f) red[f][threadIdx.x] = acc[f];
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            const int o = threadIdx.x + off;
            red[0][threadIdx.x] += red[0][o];
            red[1][threadIdx.x] = wmax(red[1][threadIdx.x], red[1][o]);
#pragma unroll
            for (int f = 2; f < 12; f += 2) {
                red[f][threadIdx.x] = wmin(red[f][threadIdx.x], red[f][o]);
                red[f + 1][threadIdx.x] = wmax(red[f + 1][threadIdx.x], red[f + 1][o]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 12) out[threadIdx.x] = red[threadIdx.x][0];
}

// Folds the per-workgroup partials in a fixed order (thread t takes
// partials t, t + 256, ..., then a tree over the threads), so the result is
// the same on every run for a given particle count.
__device__ __forceinline__ void stats_final(uint32_t n, uint32_t blocks, const float* __restrict__ partials,
                                            const uint32_t* __restrict__ counters,
                                            const uint32_t* __restrict__ clamps, float* __restrict__ stats);

__global__ __launch_bounds__(kBlock) void k_stats_final(uint32_t n, uint32_t blocks, const float* __restrict__ partials,
                                                        const uint32_t* __restrict__ counters,
                                                        const uint32_t* __restrict__ clamps,
                                                        float* __restrict__ stats) {
    stats_final(n, blocks, partials, counters, clamps, stats);
}

// One workgroup a scene: that scene's partials, counters and stats row.
__global__ __launch_bounds__(kBlock) void k_stats_final_many(SceneTable sc, const float* __restrict__ partials,
                                                             const uint32_t* __restrict__ counters,
                                                             const uint32_t* __restrict__ clamps,
                                                             float* __restrict__ stats) {
    const SceneK& k = sc.ks[blockIdx.x];
    const uint32_t n = k.slot_end - k.slot_begin;
    if (n == 0) {  // an empty scene reports zeros, as a lone sim without particles does
        if (threadIdx.x < 16) stats[(size_t)blockIdx.x * 16 + threadIdx.x] = 0.0f;
        return;
    }
    stats_final(n, stats_block_count(n), partials + (size_t)blockIdx.x * kStatsMaxBlocks * 16, counters + 2 * blockIdx.x,
                clamps + (size_t)blockIdx.x * kClampWords, stats + (size_t)blockIdx.x * 16);
}

__device__ __forceinline__ void stats_final(uint32_t n, uint32_t blocks, const float* __restrict__ partials,
                                            const uint32_t* __restrict__ counters,
                                            const uint32_t* __restrict__ clamps, float* __restrict__ stats) {
    float acc[12] = {0.0f, 0.0f, HUGE_F, -HUGE_F, HUGE_F, -HUGE_F, HUGE_F, -HUGE_F, HUGE_F, -HUGE_F, HUGE_F, -HUGE_F};
    for (uint32_t b = threadIdx.x; b < blocks; b += kBlock) {
        const float* p = partials + (size_t)b * 16;
        acc[0] += p[0];
        acc[1] = wmax(acc[1], p[1]);
#pragma unroll
        for (int f = 2; f < 12; f += 2) {
            acc[f] = wmin(acc[f], p[f]);
            acc[f + 1] = wmax(acc[f + 1], p[f + 1]);
        }
    }
    __shared__ float red[12][kBlock];
#pragma unroll
    for (int f = 0; f < 12; ++f) red[f][threadIdx.x] = acc[f];
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            const int o = threadIdx.x + off;
            red[0][threadIdx.x] += red[0][o];
            red[1][threadIdx.x] = wmax(red[1][threadIdx.x], red[1][o]);
#pragma unroll
            for (int f = 2; f < 12; f += 2) {
                red[f][threadIdx.x] = wmin(red[f][threadIdx.x], red[f][o]);
                red[f + 1][threadIdx.x] = wmax(red[f + 1][threadIdx.x], red[f + 1][o]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    stats[0] = red[0][0] / (float)n;
    stats[1] = wmax(red[1][0], 0.0f);
    stats[2] = red[2][0];
    stats[3] = red[3][0];
    stats[4] = red[8][0];
    stats[5] = red[9][0];
    stats[6] = red[5][0];
    stats[7] = red[10][0];
    stats[8] = red[11][0];
    uint32_t clamped = 0;
    for (uint32_t w = 0; w < kClampSlots; ++w) clamped += clamps[w * kClampStride];
    stats[9] = (float)clamped;
    stats[10] = (float)counters[1];
    stats[11] = red[4][0];
    stats[12] = red[6][0];
    stats[13] = red[7][0];
}

// ---- tracers -------------------------------------------------------------

__device__ __forceinline__ uint32_t pcg(uint32_t x) {
    uint32_t h = x * 747796405u + 2891336453u;
    h = ((h >> ((h >> 28u) + 4u)) ^ h) * 277803737u;
    return (h >> 22u) ^ h;
}

__device__ __forceinline__ uint32_t unorm16(float v) {
    const float c = wmin(wmax(v, 0.0f), 1.0f);
    return (uint32_t)(0.5f + 65535.0f * c);
}

// IEEE binary16 bits, round to nearest even (pack2x16float).
__device__ __forceinline__ uint32_t half_bits(float f) {
    const uint32_t x = __float_as_uint(f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u);
    if (ax >= 0x477ff000u) return sign | 0x7c00u;
    if (ax < 0x38800000u) {
        if (ax < 0x33000000u) return sign;
        const uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
        const uint32_t shift = 126u - (ax >> 23);
        uint32_t q = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1u);
        const uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (q & 1u))) q++;
        return sign | q;
    }
    uint32_t q = (ax - 0x38000000u) >> 13;
    const uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
    return sign | q;
}

__device__ __forceinline__ float half_value(uint32_t h) {
    const uint32_t sign = (h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0) {
        const float v = (float)m * (1.0f / 16777216.0f);
        return __uint_as_float(__float_as_uint(v) | sign);
    }
    if (e == 31) return __uint_as_float(sign | 0x7f800000u | (m << 13));
    return __uint_as_float(sign | ((e + 112u) << 23) | (m << 13));
}

__device__ __forceinline__ float box_extent(const SceneK& k, int a) { return -(k.box_min[a] + k.cell); }

__device__ __forceinline__ void store_tracer(const SceneK& k, uint32_t* tr, uint32_t i, float x, float y, float z,
                                             float sp) {
    const uint32_t ux = unorm16(x / box_extent(k, 0) * 0.5f + 0.5f);
    const uint32_t uy = unorm16(y / box_extent(k, 1) * 0.5f + 0.5f);
    uint2 out;
    out.x = ux | (uy << 16);
    out.y = half_bits(z) | (half_bits(sp) << 16);
    reinterpret_cast<uint2*>(tr)[i] = out;
}

__device__ __forceinline__ void respawn(const SceneK& k, uint32_t n, const float4* __restrict__ pos,
                                        const float4* __restrict__ vel, uint32_t* tr, uint32_t slot, uint32_t r) {
    const uint32_t r1 = pcg(r), r2 = pcg(r1), r3 = pcg(r2), r4 = pcg(r3);
    const float hc = k.cell * 0.5f;
    const float jx = ((float)(r2 >> 20) / 4096.0f - 0.5f) * hc;
    const float jy = ((float)(r3 >> 20) / 4096.0f - 0.5f) * hc;
    const float jz = ((float)(r4 >> 20) / 4096.0f - 0.5f) * hc;
    const uint32_t j = r1 % n;
    const float4 pj = pos[j];
    const float4 vj = vel[j];
    float q[3] = {pj.x + jx, pj.y + jy, pj.z + jz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float e = box_extent(k, a);
        q[a] = wmin(wmax(q[a], -e), e);
    }
    store_tracer(k, tr, slot, q[0], q[1], q[2], sqrtf(vj.x * vj.x + vj.y * vj.y + vj.z * vj.z));
}

// Particles of one cell sit next to each other in the resting set (it is in
// the last substep's cell order), so each wave first sums its runs of equal
// cells (a segmented scan over run ids) and the last lane of a run adds the
// run's sums: one set of atomics per run instead of per particle. Integer
// sums, so the grid holds the same values in any order.
__global__ __launch_bounds__(kBlock) void k_splat(GridGeom g, uint32_t n, const float4* __restrict__ pos,
                                                  const float4* __restrict__ vel, int32_t* __restrict__ grid) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < n;
    uint32_t c = 0xffffffffu;
    int32_t v[4] = {0, 0, 0, 0};
    if (live) {
        const float4 p = pos[i];
        c = (cell_coord(p.z, g.box_min[2], g.cell, g.dims[2]) * g.dims[1] +
             cell_coord(p.y, g.box_min[1], g.cell, g.dims[1])) *
                g.dims[0] +
            cell_coord(p.x, g.box_min[0], g.cell, g.dims[0]);
        const float4 q = vel[i];
        v[0] = sat_i32(q.x * FIXED);
        v[1] = sat_i32(q.y * FIXED);
        v[2] = sat_i32(q.z * FIXED);
        v[3] = 1;
    }
    const uint32_t lane = __lane_id();
    const uint32_t prev = (uint32_t)__shfl_up((int)c, 1);
    const bool head = lane == 0 || prev != c;
    const uint64_t heads = __ballot(head);
    const uint32_t run = (uint32_t)__popcll(heads & ((2ull << lane) - 1ull));  // run id, 1-based
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t other = (uint32_t)__shfl_up((int)run, d);
        int32_t t[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) t[a] = __shfl_up(v[a], d);
        if (lane >= d && other == run) {
#pragma unroll
            for (int a = 0; a < 4; ++a) v[a] = (int32_t)((uint32_t)v[a] + (uint32_t)t[a]);
        }
    }
    const uint32_t next_run = (uint32_t)__shfl_down((int)run, 1);
    const bool tail = lane == 63 || next_run != run;
    if (live && tail) {
        const uint32_t base = c * 4u;
#pragma unroll
        for (int a = 0; a < 4; ++a) atomicAdd(&grid[base + a], v[a]);
    }
}

__global__ __launch_bounds__(kBlock) void k_resolve(uint32_t cells, int4* __restrict__ grid, int4* __restrict__ flat) {
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= cells) return;
    flat[c] = grid[c];
    grid[c] = make_int4(0, 0, 0, 0);
}

__global__ __launch_bounds__(kBlock) void k_advect(SceneK k, StepK s, uint32_t n, const float4* __restrict__ pos,
                                                   const float4* __restrict__ vel, const int4* __restrict__ flat,
                                                   uint32_t* __restrict__ tr, uint32_t ntr) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ntr) return;
    const uint32_t r0 = pcg(i ^ pcg(s.seed));
    if ((float)r0 / 4294967295.0f < s.dt / TAU) {
        respawn(k, n, pos, vel, tr, i, r0);
        return;
    }
    const uint2 t = reinterpret_cast<const uint2*>(tr)[i];
    const float ex = box_extent(k, 0), ey = box_extent(k, 1), ez = box_extent(k, 2);
    const float px = ((float)(t.x & 0xffffu) / 65535.0f * 2.0f - 1.0f) * ex;
    const float py = ((float)(t.x >> 16) / 65535.0f * 2.0f - 1.0f) * ey;
    const float pz = half_value(t.y & 0xffffu);
    const float ow = half_value(t.y >> 16);
    const float pp[3] = {px, py, pz};
    int32_t base[3];
    float f[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float gp = (pp[a] - k.box_min[a]) / k.cell - 0.5f;
        int32_t b = sat_i32(floorf(gp));
        const int32_t hi = (int32_t)k.dims[a] - 2;
        b = b < 0 ? 0 : (b > hi ? hi : b);
        base[a] = b;
        f[a] = wmin(wmax(gp - (float)(uint32_t)b, 0.0f), 1.0f);
    }
    float mx = 0.0f, my = 0.0f, mz = 0.0f, wc = 0.0f;
#pragma unroll
    for (uint32_t kk = 0; kk < 8; ++kk) {
        const uint32_t o0 = kk & 1u, o1 = (kk >> 1) & 1u, o2 = (kk >> 2) & 1u;
        const float w0 = wmix(1.0f - f[0], f[0], (float)o0);
        const float w1 = wmix(1.0f - f[1], f[1], (float)o1);
        const float w2 = wmix(1.0f - f[2], f[2], (float)o2);
        const float weight = w0 * w1 * w2;
        const uint32_t c = (((uint32_t)base[2] + o2) * k.dims[1] + ((uint32_t)base[1] + o1)) * k.dims[0] +
                           ((uint32_t)base[0] + o0);
        const int4 cv = flat[c];
        mx += weight * (float)cv.x;
        my += weight * (float)cv.y;
        mz += weight * (float)cv.z;
        wc += weight * (float)cv.w;
    }
    float vx = 0.0f, vy = 0.0f, vz = 0.0f;
    if (wc > 0.0f) {
        const float den = wc * FIXED;
        vx = mx / den;
        vy = my / den;
        vz = mz / den;
    } else if ((float)pcg(r0 ^ 0x85ebca6bu) / 4294967295.0f < s.dt / TAU_STRAY) {
        respawn(k, n, pos, vel, tr, i, pcg(r0 ^ 0x632be59bu));
        return;
    }
    const float q0 = wmin(wmax(px + vx * s.dt, -ex), ex);
    const float q1 = wmin(wmax(py + vy * s.dt, -ey), ey);
    const float q2 = wmin(wmax(pz + vz * s.dt, -ez), ez);
    const float sp = wmax(sqrtf(vx * vx + vy * vy + vz * vz), ow * s.decay);
    store_tracer(k, tr, i, q0, q1, q2, sp);
}

}  // namespace

uint32_t stats_blocks(uint32_t n) { return stats_block_count(n); }

void launch_reduce_stats(hipStream_t st, float rho0, uint32_t n, const float* density, const Records& rest,
                         const uint32_t* counters, const uint32_t* clamps, const StatsBufs& b, KernelTimer* timer) {
    if (timer) timer->begin(st, KERNEL_REDUCE_STATS);
    hipLaunchKernelGGL(k_stats_partial, dim3(b.blocks), dim3(kBlock), 0, st, rho0, n, density, rest.vel,
                       rest.prev_vel, rest.temperature, b.partials);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(kBlock), 0, st, n, b.blocks, b.partials, counters, clamps,
                       b.stats);
    if (timer) timer->end(st, KERNEL_REDUCE_STATS);
}

void launch_reduce_stats(hipStream_t st, const SceneTable& sc, uint32_t scenes, const float* density,
                         const Records& rest, const uint32_t* counters, const uint32_t* clamps, const StatsBufs& b,
                         KernelTimer* timer) {
    if (scenes == 0) return;
    if (timer) timer->begin(st, KERNEL_REDUCE_STATS);
    hipLaunchKernelGGL(k_stats_partial_many, dim3(b.blocks, scenes), dim3(kBlock), 0, st, sc, density, rest.vel,
                       rest.prev_vel, rest.temperature, b.partials);
    hipLaunchKernelGGL(k_stats_final_many, dim3(scenes), dim3(kBlock), 0, st, sc, b.partials, counters, clamps,
                       b.stats);
    if (timer) timer->end(st, KERNEL_REDUCE_STATS);
}

void launch_tracers(hipStream_t st, const SceneK& k, const StepK& s, uint32_t n, const Records& rest,
                    const TracerBufs& b, KernelTimer* timer) {
    const uint32_t cells = k.dims[0] * k.dims[1] * k.dims[2];
    GridGeom g;
    for (int a = 0; a < 3; ++a) {
        g.box_min[a] = k.box_min[a];
        g.dims[a] = k.dims[a];
    }
    g.cell = k.cell;
    if (timer) timer->begin(st, KERNEL_TRACERS);
    if (n > 0)
        hipLaunchKernelGGL(k_splat, dim3(blocks_for(n)), dim3(kBlock), 0, st, g, n, rest.pos, rest.vel, b.vel_grid);
    hipLaunchKernelGGL(k_resolve, dim3(blocks_for(cells)), dim3(kBlock), 0, st, cells,
                       reinterpret_cast<int4*>(b.vel_grid), reinterpret_cast<int4*>(b.vel_flat));
    if (n > 0 && b.count > 0)
        hipLaunchKernelGGL(k_advect, dim3(blocks_for(b.count)), dim3(kBlock), 0, st, k, s, n, rest.pos, rest.vel,
                           reinterpret_cast<const int4*>(b.vel_flat), b.tracers, b.count);
    if (timer) timer->end(st, KERNEL_TRACERS);
}

}  // namespace fluid
