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
// Layout: every array is in slot (cell) order after the grid build. A
// workgroup owns a tile of kTile consecutive slots. All neighbours of a
// tile lie in nine slot ranges, one per stencil row (dy, dz): in cell-id
// space, row (dy, dz) of every particle of the tile falls in
// [first cell + off(dy, dz) - 1, last cell + off(dy, dz) + 1]. k_tile_regions
// records those ranges once per substep; each sweep then stages the fields it
// gathers for them into LDS with independent coalesced loads, and reads
// neighbours from LDS. A tile whose ranges exceed the LDS capacity gathers
// from global memory; the values read, and so the results, are the same.
//
// density_div: one thread a particle. It walks the stencil rows in the
// reference's cell order (x fastest, then y, then z; in-cell slot order) —
// the reference's LANES = 1 reading — testing distances only and appending
// accepted slots to the neighbour list with per-row counts; then it runs the
// per-pair math over that compact list. The list holds, per pair, the slot
// and the scalar dW/r (neighbour-major, so entry m of consecutive particles is
// contiguous); the later sweeps rebuild each gradient as dW/r times x_i - x_j,
// the product grad_q itself forms, so it is bit-identical.
//
// The list sweeps run kLanes lanes a particle, as the reference runs 8: lane l
// takes list entries l, l + kLanes, ..., and the partials fold with the
// reference's xor butterfly (lane_fold, sim_solve.wgsl:398-402). Each lane
// loads its first kPrefetch entries up front, so their latencies overlap.
#include "fluid_common.hpp"
#include "fluid_kernels.hpp"

namespace fluid {

namespace {

constexpr uint32_t kTile = 256;                      // particles a workgroup
constexpr uint32_t kLanes = 4;                       // lanes a particle in the list sweeps
constexpr uint32_t kSweepThreads = kTile * kLanes;  // 1024
constexpr uint32_t kRows = 9;
constexpr uint32_t kPrefetch = 8;                    // list entries a lane loads ahead

__device__ __forceinline__ uint32_t tile_index() { return xcd_tile(blockIdx.x, gridDim.x); }

__device__ __forceinline__ uint32_t cell_of_pos(const SceneK& k, float4 p) {
    const uint32_t cx = cell_coord(p.x, k.box_min[0], k.cell, k.dims[0]);
    const uint32_t cy = cell_coord(p.y, k.box_min[1], k.cell, k.dims[1]);
    const uint32_t cz = cell_coord(p.z, k.box_min[2], k.cell, k.dims[2]);
    return (cz * k.dims[1] + cy) * k.dims[0] + cx;
}

// One thread a tile: the tile's nine row ranges.
__global__ __launch_bounds__(kBlock) void k_tile_regions(SceneK k, uint32_t n, const uint32_t* __restrict__ starts,
                                                         const float4* __restrict__ pos, TileRegion* __restrict__ out,
                                                         uint32_t tiles) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= tiles) return;
    const uint32_t s0 = t * kTile;
    const uint32_t s1 = s0 + kTile < n ? s0 + kTile : n;
    const int64_t c_first = cell_of_pos(k, pos[s0]);
    const int64_t c_last = cell_of_pos(k, pos[s1 - 1]);
    const int64_t cells = (int64_t)k.dims[0] * k.dims[1] * k.dims[2];
    const int64_t plane = (int64_t)k.dims[0] * k.dims[1];
    TileRegion R;
    uint32_t total = 0;
#pragma unroll
    for (uint32_t r = 0; r < kRows; ++r) {
        const int64_t dz = (int64_t)(r / 3) - 1, dy = (int64_t)(r % 3) - 1;
        const int64_t off = dz * plane + dy * (int64_t)k.dims[0];
        int64_t lo = c_first + off - 1, hi = c_last + off + 1;
        if (lo < 0) lo = 0;
        if (hi > cells - 1) hi = cells - 1;
        uint32_t a = 0, e = 0;
        if (lo <= hi) {
            a = starts[lo];
            e = starts[hi + 1];
        }
        R.delta[r] = (int32_t)total - (int32_t)a;
        R.base[r] = total;
        total += e - a;
    }
    R.total = total;
    R.pad = 0;
    out[t] = R;
}

// A tile's region as uniform scalars (SGPRs): the staging loop selects the
// row of each flat LDS index among them.
struct RegionS {
    int32_t d0, d1, d2, d3, d4, d5, d6, d7, d8;
    uint32_t b1, b2, b3, b4, b5, b6, b7, b8;
    uint32_t total;
};

__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

__device__ __forceinline__ RegionS load_region(const TileRegion* __restrict__ regions, uint32_t tile) {
    const TileRegion* R = regions + tile;
    RegionS s;
    s.d0 = (int32_t)uni((uint32_t)R->delta[0]);
    s.d1 = (int32_t)uni((uint32_t)R->delta[1]);
    s.d2 = (int32_t)uni((uint32_t)R->delta[2]);
    s.d3 = (int32_t)uni((uint32_t)R->delta[3]);
    s.d4 = (int32_t)uni((uint32_t)R->delta[4]);
    s.d5 = (int32_t)uni((uint32_t)R->delta[5]);
    s.d6 = (int32_t)uni((uint32_t)R->delta[6]);
    s.d7 = (int32_t)uni((uint32_t)R->delta[7]);
    s.d8 = (int32_t)uni((uint32_t)R->delta[8]);
    s.b1 = uni(R->base[1]);
    s.b2 = uni(R->base[2]);
    s.b3 = uni(R->base[3]);
    s.b4 = uni(R->base[4]);
    s.b5 = uni(R->base[5]);
    s.b6 = uni(R->base[6]);
    s.b7 = uni(R->base[7]);
    s.b8 = uni(R->base[8]);
    s.total = uni(R->total);
    return s;
}

// The delta of the region row holding flat LDS index idx: that of the last
// row whose base <= idx (empty rows share the next row's base, so the last one
// is the non-empty one).
// Accumulated as d0 plus the step at each row boundary passed (a chain of
// adds, which keeps every value in registers).
__device__ __forceinline__ int32_t region_delta(const RegionS& R, uint32_t idx) {
    int32_t d = R.d0;
    d += (int32_t)(idx >= R.b1) * (R.d1 - R.d0);
    d += (int32_t)(idx >= R.b2) * (R.d2 - R.d1);
    d += (int32_t)(idx >= R.b3) * (R.d3 - R.d2);
    d += (int32_t)(idx >= R.b4) * (R.d4 - R.d3);
    d += (int32_t)(idx >= R.b5) * (R.d5 - R.d4);
    d += (int32_t)(idx >= R.b6) * (R.d6 - R.d5);
    d += (int32_t)(idx >= R.b7) * (R.d7 - R.d6);
    d += (int32_t)(idx >= R.b8) * (R.d8 - R.d7);
    return d;
}

// Every sweep keeps the nine row deltas in LDS for its lanes' row lookups.
__device__ __forceinline__ void load_deltas(const TileRegion* __restrict__ regions, uint32_t tile, int32_t* s_delta) {
    if (threadIdx.x < kRows) s_delta[threadIdx.x] = regions[tile].delta[threadIdx.x];
}

__device__ __forceinline__ uint32_t row_count(uint64_t rows, uint32_t r) { return (uint32_t)(rows >> (7u * r)) & 127u; }

// Advances (r, row_end) until entry m lies in row r.
__device__ __forceinline__ void seek_row(uint64_t rows, uint32_t m, uint32_t& r, uint32_t& row_end) {
    while (m >= row_end && r + 1 < kRows) {
        ++r;
        row_end += row_count(rows, r);
    }
}

// lane_fold over kLanes = 4 consecutive lanes: xor 2, then xor 1 (lane 0
// holds (p0 + p2) + (p1 + p3)).
__device__ __forceinline__ float fold4(float v) {
    const float t = v + __shfl_xor(v, 2);
    return t + __shfl_xor(t, 1);
}
__device__ __forceinline__ f3 fold4(f3 v) { return mk(fold4(v.x), fold4(v.y), fold4(v.z)); }

__device__ __forceinline__ f3 listed_grad(f3 pi, f3 pj, float g) { return scl(g, sub(pi, pj)); }

// Copies `total` staged slots into LDS: NT threads, ITERS slots each, all
// loads issued before the stores. `load(slot, it)` fetches into registers,
// `store(idx, it)` writes them to LDS.
template <uint32_t NT, uint32_t ITERS, class Load, class Store>
__device__ __forceinline__ void stage_flat(const RegionS& R, Load&& load, Store&& store) {
    for (uint32_t chunk = 0; chunk < R.total; chunk += NT * ITERS) {
#pragma unroll
        for (uint32_t it = 0; it < ITERS; ++it) {
            const uint32_t idx = chunk + it * NT + threadIdx.x;
            if (idx < R.total) load((uint32_t)((int32_t)idx - region_delta(R, idx)), it);
        }
#pragma unroll
        for (uint32_t it = 0; it < ITERS; ++it) {
            const uint32_t idx = chunk + it * NT + threadIdx.x;
            if (idx < R.total) store(idx, it);
        }
    }
}

// ---- density_div -------------------------------------------------------------

struct DensitySums {
    float w_sum, rn, gsq, drho;
    f3 gsum;
};

__device__ __forceinline__ float density_pair(const SceneK& k, f3 pi, f3 vi, f3 pj, f3 vj, DensitySums& a) {
    const f3 x = sub(pi, pj);
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
    a.drho += dot3(sub(vi, vj), gw);
    return g;
}

// Distances from the particle to the faces of its own cell: a stencil row or
// end cell farther than the support holds no neighbour, so the walk skips it
// (the accepted neighbours and their order are those of the full walk).
struct Reach {
    float dl[3], dh[3];
    float lim;
};

__device__ __forceinline__ Reach make_reach(const SceneK& k, f3 p, uint32_t bx, uint32_t by, uint32_t bz) {
    Reach st;
    const float pp[3] = {p.x, p.y, p.z};
    const uint32_t b[3] = {bx, by, bz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float lo = k.box_min[a] + (float)b[a] * k.cell;
        const float hi = k.box_min[a] + (float)(b[a] + 1u) * k.cell;
        st.dl[a] = wmax(pp[a] - lo, 0.0f);
        st.dh[a] = wmax(hi - pp[a], 0.0f);
    }
    st.lim = k.support_sq * 1.001f;  // a margin over the rounding of the faces
    return st;
}

// LDS layout of density_div and den_kappa: per slot the position and the
// velocity, 24 bytes.
struct PvLds {
    float4* pv;  // x, y, z, vx
    float2* vw;  // vy, vz
};

template <bool STAGED>
__device__ __forceinline__ void load_pv(const PvLds& L, const float4* __restrict__ pos,
                                        const float4* __restrict__ vel, uint32_t j, int32_t delta, f3& p, f3& v) {
    if (STAGED) {
        const float4 a = L.pv[j + delta];
        const float2 b = L.vw[j + delta];
        p = mk(a.x, a.y, a.z);
        v = mk(a.w, b.x, b.y);
    } else {
        p = xyz(pos[j]);
        v = xyz(vel[j]);
    }
}

template <uint32_t NT>
__device__ __forceinline__ void stage_pv(const RegionS& R, const float4* __restrict__ pos,
                                         const float4* __restrict__ vel, const PvLds& L) {
    constexpr uint32_t IT = 4;
    float4 pp[IT], vv[IT];
    stage_flat<NT, IT>(
        R,
        [&](uint32_t slot, uint32_t it) {
            pp[it] = pos[slot];
            vv[it] = vel[slot];
        },
        [&](uint32_t idx, uint32_t it) {
            L.pv[idx] = make_float4(pp[it].x, pp[it].y, pp[it].z, vv[it].x);
            L.vw[idx] = make_float2(vv[it].y, vv[it].z);
        });
}

template <bool STAGED>
__device__ void density_body(const SceneK& k, const StepK& s, uint32_t i, const uint32_t* __restrict__ starts,
                             const float4* __restrict__ pos, const float4* __restrict__ vel, const SolveBufs& b,
                             const int32_t* deltas, const PvLds& L) {
    const f3 pi = xyz(pos[i]);
    const f3 vi = xyz(vel[i]);
    const uint32_t bx = cell_coord(pi.x, k.box_min[0], k.cell, k.dims[0]);
    const uint32_t by = cell_coord(pi.y, k.box_min[1], k.cell, k.dims[1]);
    const uint32_t bz = cell_coord(pi.z, k.box_min[2], k.cell, k.dims[2]);
    const Reach st = make_reach(k, pi, bx, by, bz);
    uint2* __restrict__ list = b.nbr + i;
    const uint32_t stride = b.stride, cap = b.cap;
    // Phase 1: distance tests only, appending accepted slots and row counts.
    uint32_t found = 0;
    uint64_t rows = 0;
#pragma unroll 1
    for (uint32_t r = 0; r < kRows; ++r) {
        const int dz = (int)(r / 3) - 1, dy = (int)(r % 3) - 1;
        const int cz = (int)bz + dz, cy = (int)by + dy;
        if (cz < 0 || cz >= (int)k.dims[2] || cy < 0 || cy >= (int)k.dims[1]) continue;
        const float ez = dz < 0 ? st.dl[2] : (dz > 0 ? st.dh[2] : 0.0f);
        const float ey = dy < 0 ? st.dl[1] : (dy > 0 ? st.dh[1] : 0.0f);
        const float e2 = ey * ey + ez * ez;
        if (e2 > st.lim) continue;
        const uint32_t x0 = (bx > 0 && st.dl[0] * st.dl[0] + e2 <= st.lim) ? bx - 1 : bx;
        const uint32_t x1 = (bx + 1 < k.dims[0] && st.dh[0] * st.dh[0] + e2 <= st.lim) ? bx + 1 : bx;
        const uint32_t row = ((uint32_t)cz * k.dims[1] + (uint32_t)cy) * k.dims[0];
        const uint32_t k0 = starts[row + x0], k1 = starts[row + x1 + 1];
        const int32_t delta = deltas[r];
        uint32_t in_row = 0;
#pragma unroll 4
        for (uint32_t kk = k0; kk < k1; ++kk) {
            f3 pj;
            if (STAGED) {
                const float4 a = L.pv[kk + delta];
                pj = mk(a.x, a.y, a.z);
            } else {
                pj = xyz(pos[kk]);
            }
            const f3 x = sub(pi, pj);
            const float r2 = dot3(x, x);
            if (r2 < k.support_sq && kk != i) {
                if (found < cap) {
                    list[(size_t)found * stride].x = kk;
                    in_row++;
                }
                found++;
            }
        }
        rows |= (uint64_t)in_row << (7u * r);
    }
    // Phase 2: the sums over the compact list, in list order, in chunks of
    // kPrefetch entries whose slots are read back together.
    DensitySums a{k.sigma, 0.0f, 0.0f, 0.0f, mk(0.0f, 0.0f, 0.0f)};
    const uint32_t nn = found < cap ? found : cap;
    uint32_t r = 0, row_end = row_count(rows, 0);
    for (uint32_t m0 = 0; m0 < nn; m0 += kPrefetch) {
        uint32_t js[kPrefetch];
#pragma unroll
        for (uint32_t t = 0; t < kPrefetch; ++t)
            js[t] = m0 + t < nn ? list[(size_t)(m0 + t) * stride].x : 0u;
#pragma unroll
        for (uint32_t t = 0; t < kPrefetch; ++t) {
            const uint32_t m = m0 + t;
            if (m < nn) {
                seek_row(rows, m, r, row_end);
                f3 pj, vj;
                load_pv<STAGED>(L, pos, vel, js[t], deltas[r], pj, vj);
                const float g = density_pair(k, pi, vi, pj, vj, a);
                list[(size_t)m * stride].y = __float_as_uint(g);
            }
        }
    }
    if (found > cap) {
        // Past the cap the list keeps its first `cap` neighbours (the later
        // sweeps see only those, as in the reference) but the density sums
        // take every neighbour: walk again for the ones the list dropped.
        atomicAdd(&b.counters[1], found - cap);
        uint32_t seen = 0;
        for (uint32_t rr = 0; rr < kRows; ++rr) {
            const int dz = (int)(rr / 3) - 1, dy = (int)(rr % 3) - 1;
            const int cz = (int)bz + dz, cy = (int)by + dy;
            if (cz < 0 || cz >= (int)k.dims[2] || cy < 0 || cy >= (int)k.dims[1]) continue;
            const uint32_t x0 = bx > 0 ? bx - 1 : bx;
            const uint32_t x1 = bx + 1 < k.dims[0] ? bx + 1 : bx;
            const uint32_t row = ((uint32_t)cz * k.dims[1] + (uint32_t)cy) * k.dims[0];
            const uint32_t k0 = starts[row + x0], k1 = starts[row + x1 + 1];
            for (uint32_t kk = k0; kk < k1; ++kk) {
                const f3 pj = xyz(pos[kk]);
                const f3 x = sub(pi, pj);
                if (dot3(x, x) < k.support_sq && kk != i) {
                    if (seen >= cap) density_pair(k, pi, vi, pj, xyz(vel[kk]), a);
                    seen++;
                }
            }
        }
    }
    b.nbr_n[i] = nn;
    b.nbr_rows[i] = rows;
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

__global__ __launch_bounds__(kTile) void k_density_div(SceneK k, StepK s, uint32_t n, const uint32_t* __restrict__ starts,
                                                       const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                       SolveBufs b) {
    extern __shared__ float4 smem[];
    __shared__ int32_t s_delta[kRows];
    const uint32_t tile = tile_index();
    const RegionS R = load_region(b.regions, tile);
    const bool staged = R.total <= b.cap_density;
    const PvLds L{smem, reinterpret_cast<float2*>(smem + b.cap_density)};
    load_deltas(b.regions, tile, s_delta);
    if (staged) stage_pv<kTile>(R, pos, vel, L);
    __syncthreads();
    const uint32_t i = tile * kTile + threadIdx.x;
    if (i >= n) return;
    if (staged)
        density_body<true>(k, s, i, starts, pos, vel, b, s_delta, L);
    else
        density_body<false>(k, s, i, starts, pos, vel, b, s_delta, L);
}

// ---- list sweeps -------------------------------------------------------------

struct LaneId {
    uint32_t tile, p, lane;
    bool live;
};

__device__ __forceinline__ LaneId lane_id(uint32_t n) {
    LaneId id;
    id.tile = tile_index();
    id.p = id.tile * kTile + threadIdx.x / kLanes;
    id.lane = threadIdx.x % kLanes;
    id.live = id.p < n;
    return id;
}

// Walks lane `lane`'s share of particle p's list in order, calling
// f(slot, dW/r, delta of the slot's row) for each entry.
template <class F>
__device__ __forceinline__ void for_list(const SolveBufs& b, const int32_t* deltas, uint32_t p, uint32_t lane, F&& f) {
    const uint32_t nn = b.nbr_n[p];
    const uint64_t rows = b.nbr_rows[p];
    const uint2* __restrict__ list = b.nbr + p;
    uint32_t r = 0, row_end = row_count(rows, 0);
    uint2 e[kPrefetch];
#pragma unroll
    for (uint32_t t = 0; t < kPrefetch; ++t) {
        const uint32_t m = lane + t * kLanes;
        e[t] = m < nn ? list[(size_t)m * b.stride] : make_uint2(0u, 0u);
    }
#pragma unroll
    for (uint32_t t = 0; t < kPrefetch; ++t) {
        const uint32_t m = lane + t * kLanes;
        if (m < nn) {
            seek_row(rows, m, r, row_end);
            f(e[t].x, __uint_as_float(e[t].y), deltas[r]);
        }
    }
    for (uint32_t m = lane + kPrefetch * kLanes; m < nn; m += kLanes) {
        seek_row(rows, m, r, row_end);
        const uint2 q = list[(size_t)m * b.stride];
        f(q.x, __uint_as_float(q.y), deltas[r]);
    }
}

// Stages position and kd (the gathers of div_apply, den_apply and
// forces_den_apply) as one float4 per slot.
__device__ __forceinline__ void stage_pos_kd(const RegionS& R, const float4* __restrict__ pos,
                                             const float* __restrict__ kd, float4* lds) {
    constexpr uint32_t IT = 4;
    float4 pp[IT];
    float kk[IT];
    stage_flat<kSweepThreads, IT>(
        R,
        [&](uint32_t slot, uint32_t it) {
            pp[it] = pos[slot];
            kk[it] = kd[slot];
        },
        [&](uint32_t idx, uint32_t it) { lds[idx] = make_float4(pp[it].x, pp[it].y, pp[it].z, kk[it]); });
}

// Lane `lane` of particle p: the partial sum of (kd_i + kd_j) grad W_ij over
// its share of the list (kappa_partial, sim_solve.wgsl:596-606).
template <bool STAGED>
__device__ __forceinline__ f3 kappa_partial(const SolveBufs& b, const int32_t* deltas, uint32_t p, uint32_t lane,
                                            f3 pi, float kdi, const float4* __restrict__ pos, const float4* lds) {
    f3 dv = mk(0.0f, 0.0f, 0.0f);
    for_list(b, deltas, p, lane, [&](uint32_t j, float g, int32_t delta) {
        float4 q;
        if (STAGED) {
            q = lds[j + delta];
        } else {
            const float4 pj = pos[j];
            q = make_float4(pj.x, pj.y, pj.z, b.kd[j]);
        }
        dv = add(dv, scl(kdi + q.w, listed_grad(pi, xyz(q), g)));
    });
    return dv;
}

// div_apply / den_apply: v -= dt (m sum (kd_i + kd_j) gradW + kd_i wall_grad).
// Reads only kd of the neighbours, so the velocity update is in place.
__global__ __launch_bounds__(kSweepThreads) void k_kappa_apply(SceneK k, StepK s, uint32_t n,
                                                               const float4* __restrict__ pos, float4* __restrict__ vel,
                                                               SolveBufs b) {
    extern __shared__ float4 smem[];
    __shared__ int32_t s_delta[kRows];
    const LaneId id = lane_id(n);
    const RegionS R = load_region(b.regions, id.tile);
    load_deltas(b.regions, id.tile, s_delta);
    const bool staged = R.total <= b.cap_kappa;
    if (staged) stage_pos_kd(R, pos, b.kd, smem);
    __syncthreads();
    f3 dv = mk(0.0f, 0.0f, 0.0f), pi = dv;
    float kdi = 0.0f;
    if (id.live) {
        pi = xyz(pos[id.p]);
        kdi = b.kd[id.p];
        dv = staged ? kappa_partial<true>(b, s_delta, id.p, id.lane, pi, kdi, pos, smem)
                    : kappa_partial<false>(b, s_delta, id.p, id.lane, pi, kdi, pos, smem);
    }
    const f3 total = fold4(dv);
    if (id.live && id.lane == 0) {
        const float4 wg = b.wall_grad[id.p];
        const f3 full = add(scl(k.mass, total), scl(kdi, xyz(wg)));
        const float4 w = vel[id.p];
        vel[id.p] = make_float4(w.x - s.dt * full.x, w.y - s.dt * full.y, w.z - s.dt * full.z, w.w);
    }
}

// forces_eval's gathers: position, velocity, density, near-density and
// temperature, 36 bytes a slot as float4 (x, y, z, vx), float4 (vy, vz,
// rho, rho_near), float (temperature).
struct ForcesLds {
    float4* a;
    float4* c;
    float* t;
};

__global__ __launch_bounds__(kSweepThreads) void k_forces_eval(SceneK k, uint32_t n, const float4* __restrict__ pos,
                                                               const float4* __restrict__ vel,
                                                               const float* __restrict__ temperature,
                                                               const float* __restrict__ prev_pressure, SolveBufs b) {
    extern __shared__ float4 smem[];
    __shared__ int32_t s_delta[kRows];
    const LaneId id = lane_id(n);
    const RegionS R = load_region(b.regions, id.tile);
    load_deltas(b.regions, id.tile, s_delta);
    const bool staged = R.total <= b.cap_forces;
    const ForcesLds L{smem, smem + b.cap_forces, reinterpret_cast<float*>(smem + 2 * b.cap_forces)};
    if (staged) {
        constexpr uint32_t IT = 2;
        float4 pp[IT], vv[IT];
        float dd[IT], nr[IT], tt[IT];
        stage_flat<kSweepThreads, IT>(
            R,
            [&](uint32_t slot, uint32_t it) {
                pp[it] = pos[slot];
                vv[it] = vel[slot];
                dd[it] = b.density[slot];
                nr[it] = b.rho_near[slot];
                tt[it] = temperature[slot];
            },
            [&](uint32_t idx, uint32_t it) {
                L.a[idx] = make_float4(pp[it].x, pp[it].y, pp[it].z, vv[it].x);
                L.c[idx] = make_float4(vv[it].y, vv[it].z, dd[it], nr[it]);
                L.t[idx] = tt[it];
            });
    }
    __syncthreads();
    f3 visc = mk(0.0f, 0.0f, 0.0f), blend = visc, nearf = visc, tension = visc, pi = visc;
    float heat = 0.0f, rho_i = 0.0f;
    if (id.live) {
        const uint32_t i = id.p;
        pi = xyz(pos[i]);
        const f3 vi = xyz(vel[i]);
        const float w_i = b.rho_near[i], temp_i = temperature[i];
        rho_i = b.density[i];
        const float rhoi = rho_i;
        for_list(b, s_delta, i, id.lane, [&](uint32_t j, float gs, int32_t delta) {
            f3 pj, vj;
            float rho_j, near_j, temp_j;
            if (staged) {
                const uint32_t q = j + delta;
                const float4 a = L.a[q], c = L.c[q];
                pj = mk(a.x, a.y, a.z);
                vj = mk(a.w, c.x, c.y);
                rho_j = c.z;
                near_j = c.w;
                temp_j = L.t[q];
            } else {
                pj = xyz(pos[j]);
                vj = xyz(vel[j]);
                rho_j = b.density[j];
                near_j = b.rho_near[j];
                temp_j = temperature[j];
            }
            const f3 x = sub(pi, pj);
            const float r2 = dot3(x, x);
            const float rr = sqrtf(r2);
            const f3 g = scl(gs, x);
            const float f = dot3(x, g) / (r2 + k.eps_h2);
            const float pair = k.mass / (rhoi * rho_j) * f;
            const f3 dv = sub(vi, vj);
            const float wn = wmax(1.0f - rr * k.inv_h, 0.0f);
            const float rm = wmax(rr, 1e-5f);
            const float sn = (w_i + near_j) * wn * wn;
            nearf = add(nearf, mk(sn * (x.x / rm), sn * (x.y / rm), sn * (x.z / rm)));
            visc = add(visc, scl(k.visc2 * pair, dv));
            blend = sub(blend, scl(k.mass / rho_j * kernel_q(k, rr * k.inv_h), dv));
            heat += k.heat_cond * pair * (temp_i - temp_j);
            heat -= k.heat_visc * pair * dot3(dv, dv);
            const float corr = 2.0f * k.rho0 / (rhoi + rho_j);
            const float cm = corr * k.mass * cohesion(rr, k.coh_c, k.coh_k, k.coh_near);
            const float rg = wmax(rr, 1e-9f);
            tension = sub(tension, mk(cm * x.x / rg, cm * x.y / rg, cm * x.z / rg));
        });
    }
    visc = fold4(visc);
    blend = fold4(blend);
    nearf = fold4(nearf);
    tension = fold4(tension);
    heat = fold4(heat);
    if (id.live && id.lane == 0) {
        const uint32_t i = id.p;
        const f3 adh = scl(k.beta * k.rho0, wall_adh_sum(k, pi));
        const f3 acc = add(add(add(visc, scl(K_NEAR, nearf)), scl(k.gamma, tension)), adh);
        b.accel[i] = make_float4(acc.x, acc.y, acc.z, heat);
        b.xsph[i] = make_float4(blend.x, blend.y, blend.z, 0.0f);
        const float half = 0.5f * prev_pressure[i];
        b.kd[i] = half / (rho_i * rho_i);
        b.pressure[i] = half;
    }
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

__global__ __launch_bounds__(kSweepThreads) void k_forces_den_apply(SceneK k, StepK s, uint32_t n,
                                                                    const float4* __restrict__ pos,
                                                                    float4* __restrict__ vel,
                                                                    float* __restrict__ temperature, SolveBufs b) {
    extern __shared__ float4 smem[];
    __shared__ int32_t s_delta[kRows];
    const LaneId id = lane_id(n);
    const RegionS R = load_region(b.regions, id.tile);
    load_deltas(b.regions, id.tile, s_delta);
    const bool staged = R.total <= b.cap_kappa;
    if (staged) stage_pos_kd(R, pos, b.kd, smem);
    __syncthreads();
    f3 dv = mk(0.0f, 0.0f, 0.0f), pi = dv;
    float kdi = 0.0f;
    if (id.live) {
        pi = xyz(pos[id.p]);
        kdi = b.kd[id.p];
        dv = staged ? kappa_partial<true>(b, s_delta, id.p, id.lane, pi, kdi, pos, smem)
                    : kappa_partial<false>(b, s_delta, id.p, id.lane, pi, kdi, pos, smem);
    }
    const f3 total = fold4(dv);
    if (id.live && id.lane == 0) {
        const uint32_t i = id.p;
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
        const f3 full = add(scl(k.mass, total), scl(kdi, xyz(wg)));
        vel[i] = make_float4(forced.x - s.dt * full.x, forced.y - s.dt * full.y, forced.z - s.dt * full.z, was.w);
    }
}

__global__ __launch_bounds__(kSweepThreads) void k_den_kappa(SceneK k, StepK s, uint32_t n,
                                                             const float4* __restrict__ pos,
                                                             const float4* __restrict__ vel, SolveBufs b) {
    extern __shared__ float4 smem[];
    __shared__ int32_t s_delta[kRows];
    const LaneId id = lane_id(n);
    const RegionS R = load_region(b.regions, id.tile);
    load_deltas(b.regions, id.tile, s_delta);
    const bool staged = R.total <= b.cap_kappa_vel;
    const PvLds L{smem, reinterpret_cast<float2*>(smem + b.cap_kappa_vel)};
    if (staged) stage_pv<kSweepThreads>(R, pos, vel, L);
    __syncthreads();
    float drho = 0.0f;
    f3 vi = mk(0.0f, 0.0f, 0.0f);
    if (id.live) {
        const uint32_t i = id.p;
        const f3 pi = xyz(pos[i]);
        vi = xyz(vel[i]);
        const f3 vii = vi;
        for_list(b, s_delta, i, id.lane, [&](uint32_t j, float g, int32_t delta) {
            f3 pj, vj;
            if (staged)
                load_pv<true>(L, pos, vel, j, delta, pj, vj);
            else
                load_pv<false>(L, pos, vel, j, 0, pj, vj);
            drho += dot3(sub(vii, vj), listed_grad(pi, pj, g));
        });
    }
    const float total = fold4(drho);
    if (id.live && id.lane == 0) {
        const uint32_t i = id.p;
        const float rho = b.density[i];
        const float4 wg = b.wall_grad[i];
        const float d = k.mass * total + dot3(vi, xyz(wg));
        const float rho_star = rho + s.dt * d;
        const float kap = wmax(rho_star - k.rho0, 0.0f) * b.alpha[i] / (s.dt * s.dt);
        b.kd[i] = kap / rho;
        b.pressure[i] = b.pressure[i] + kap * rho;
    }
}

// Reads the working set, writes the resting set at the same slot.
__global__ __launch_bounds__(kBlock) void k_integrate(SceneK k, StepK s, uint32_t n, Records work, Records rest,
                                                      SolveBufs b) {
    const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
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

constexpr size_t kPvEntry = sizeof(float4) + sizeof(float2);          // 24
constexpr size_t kKappaEntry = sizeof(float4);                         // 16
constexpr size_t kForcesEntry = 2 * sizeof(float4) + sizeof(float);    // 36

}  // namespace

uint32_t solve_tiles(uint32_t n) { return (n + kTile - 1) / kTile; }

// Sizes the sweeps' LDS staging areas from the device's per-workgroup LDS
// (less 1 KB for the kernels' static LDS) and raises the kernels' dynamic-LDS
// limit to that size. If the device refuses the larger limit, the areas fit
// the default 64 KB.
size_t configure_solve(SolveBufs& b, size_t lds_bytes) {
    size_t avail = lds_bytes > 1024 ? lds_bytes - 1024 : 0;
    const void* fns[] = {reinterpret_cast<const void*>(&k_density_div), reinterpret_cast<const void*>(&k_kappa_apply),
                         reinterpret_cast<const void*>(&k_forces_eval),
                         reinterpret_cast<const void*>(&k_forces_den_apply),
                         reinterpret_cast<const void*>(&k_den_kappa)};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)avail) != hipSuccess)
            avail = 65536 - 1024;
    (void)hipGetLastError();
    // Keep capacities multiples of 64 slots; a typical tile region holds about
    // 9 x 34 cells x 8 slots = 2,448 slots at the configurations' spacing.
    auto cap = [&](size_t entry, size_t want) {
        size_t c = avail / entry;
        if (c > want) c = want;
        return (uint32_t)(c / 64 * 64);
    };
    b.cap_density = cap(kPvEntry, 3072);
    b.cap_kappa = cap(kKappaEntry, 4096);
    b.cap_kappa_vel = cap(kPvEntry, 3072);
    b.cap_forces = cap(kForcesEntry, 3072);
    return avail;
}

static void launch_regions(hipStream_t st, const SceneK& k, uint32_t n, const uint32_t* starts, const Records& work,
                           const SolveBufs& b) {
    const uint32_t tiles = solve_tiles(n);
    hipLaunchKernelGGL(k_tile_regions, dim3(blocks_for(tiles)), dim3(kBlock), 0, st, k, n, starts, work.pos,
                       b.regions, tiles);
}

void launch_density_only(hipStream_t st, const SceneK& k, const StepK& s, uint32_t n, const uint32_t* starts,
                         const Records& work, const SolveBufs& b, KernelTimer* timer) {
    if (n == 0) return;
    if (timer) timer->begin(st, KERNEL_DENSITY_DIV);
    launch_regions(st, k, n, starts, work, b);
    hipLaunchKernelGGL(k_density_div, dim3(solve_tiles(n)), dim3(kTile), b.cap_density * kPvEntry, st, k, s, n,
                       starts, work.pos, work.vel, b);
    if (timer) timer->end(st, KERNEL_DENSITY_DIV);
}

void launch_solve(hipStream_t st, const SceneK& k, const StepK& s, uint32_t n, const uint32_t* starts,
                  const Records& work, const Records& rest, const SolveBufs& b, uint32_t refine,
                  KernelTimer* timer) {
    if (n == 0) return;
    const dim3 tiles(solve_tiles(n)), sweep(kSweepThreads);
    const size_t lds_kappa = b.cap_kappa * kKappaEntry;
    launch_density_only(st, k, s, n, starts, work, b, timer);
    if (timer) timer->begin(st, KERNEL_DIV_APPLY);
    hipLaunchKernelGGL(k_kappa_apply, tiles, sweep, lds_kappa, st, k, s, n, work.pos, work.vel, b);
    if (timer) timer->end(st, KERNEL_DIV_APPLY);
    if (timer) timer->begin(st, KERNEL_FORCES_EVAL);
    hipLaunchKernelGGL(k_forces_eval, tiles, sweep, b.cap_forces * kForcesEntry, st, k, n, work.pos, work.vel,
                       work.temperature, work.prev_pressure, b);
    if (timer) timer->end(st, KERNEL_FORCES_EVAL);
    if (timer) timer->begin(st, KERNEL_FORCES_DEN_APPLY);
    hipLaunchKernelGGL(k_forces_den_apply, tiles, sweep, lds_kappa, st, k, s, n, work.pos, work.vel,
                       work.temperature, b);
    if (timer) timer->end(st, KERNEL_FORCES_DEN_APPLY);
    for (uint32_t r = 0; r < refine; ++r) {
        if (timer) timer->begin(st, KERNEL_DEN_KAPPA);
        hipLaunchKernelGGL(k_den_kappa, tiles, sweep, b.cap_kappa_vel * kPvEntry, st, k, s, n, work.pos, work.vel,
                           b);
        if (timer) timer->end(st, KERNEL_DEN_KAPPA);
        if (timer) timer->begin(st, KERNEL_DEN_APPLY);
        hipLaunchKernelGGL(k_kappa_apply, tiles, sweep, lds_kappa, st, k, s, n, work.pos, work.vel, b);
        if (timer) timer->end(st, KERNEL_DEN_APPLY);
    }
    if (timer) timer->begin(st, KERNEL_INTEGRATE);
    hipLaunchKernelGGL(k_integrate, dim3(blocks_for(n)), dim3(kBlock), 0, st, k, s, n, work, rest, b);
    if (timer) timer->end(st, KERNEL_INTEGRATE);
}

}  // namespace fluid
