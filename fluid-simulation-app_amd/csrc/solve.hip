// The DFSPH solve for gfx950: density and the DFSPH factor, the divergence
// and constant-density pressure solves, the force terms, and integrate.
//
// Replaces the reference's sweeps (crates/fluid-core/src/sim_solve.wgsl):
//   density_div        411-518   27-cell walk; density, near-density, alpha,
//                                divergence kappa, wall gradient, neighbour list
//   div_apply          665-668   (kappa_partial + apply_kappa_wide, 594-621)
//   forces_eval        524-592   } one sweep (k_forces): both read the same list,
//   forces_den_apply   630-663   } and the warm start's kd needs no neighbour
//   den_kappa          674-698
//   den_apply          700-703
//   integrate          709-756
//
// Layout: every array is in slot (cell) order after the grid build. A
// workgroup owns a tile of kTile consecutive slots, kLanes threads a particle
// (kTile * kLanes threads). All neighbours of a tile lie in nine slot
// ranges, one per stencil row (dy, dz): in cell-id space, row (dy, dz) of
// every particle of the tile falls in [first cell + off(dy, dz) - 1, last
// cell + off(dy, dz) + 1]. k_tile_regions records those ranges once per
// substep; when they fit the staging capacity, every sweep copies the fields
// it gathers for them into LDS with independent coalesced loads, and the
// neighbour list of the tile's particles holds LDS indices. Otherwise the
// list holds slots and the sweeps gather from global memory; the values read,
// and so the results, are the same.
//
// density_div walks the stencil rows in the reference's cell order (x
// fastest, then y, then z; in-cell slot order) with the particle's four lanes
// testing consecutive candidates; a ballot over the four orders the accepted
// ones, so the list is in walk order. The list sweeps, density_div's pair sums
// included, then split the list over the lanes — lane l takes entries l, l +
// 4, ... — and fold the partials with the reference's xor butterfly
// (lane_fold, sim_solve.wgsl:398-402) over four lanes. The list holds, per
// pair, the neighbour and the scalar dW/r; the later sweeps rebuild each
// gradient as dW/r times x_i - x_j, the product grad_q itself forms, so it is
// bit-identical.
#include <chrono>

#include "fluid_common.hpp"
#include "fluid_kernels.hpp"

namespace fluid {
namespace FLUID_TILE_NS {

namespace {

constexpr uint32_t kTile = kSolveTile;              // particles a workgroup
constexpr uint32_t kLanes = kListLanes;             // lanes a particle
constexpr uint32_t kSweepThreads = kTile * kLanes;
constexpr uint32_t kRows = 9;
constexpr uint32_t kPrefetch = 4;                   // list quads a lane loads ahead
// LDS slots a staged array keeps past the staging capacity, so that a read up
// to 7 slots past any row's end stays inside the array.
constexpr uint32_t kStagePad = 8;
__host__ __device__ __forceinline__ uint32_t stage_slots(const SolveBufs& b) { return b.stage_cap + kStagePad; }

// The workgroup's tile. Consecutive list sweeps run in opposite tile orders:
// one sweep's working set (the neighbour list alone is ~200 bytes a
// particle) is about the size of the 256 MB memory-side cache, and a sweep
// that starts where the previous one ended finds its first tiles' data still
// there, where the same order every time would evict each line just before
// it is needed again.
template <class Sc>
__device__ __forceinline__ uint32_t tile_index(const Sc& sc) {
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    return sc.tile0 + (sc.flip ? gridDim.x - 1u - t : t);
}

// One thread a tile: the tile's nine row ranges (cells of the tile's own
// scene; starts indexed at the scene's cell offset).
template <class Sc>
__global__ __launch_bounds__(kBlock) void k_tile_regions(Sc sc, const uint32_t* __restrict__ starts,
                                                         const float4* __restrict__ pos, TileRegion* __restrict__ out,
                                                         uint32_t tiles, uint32_t stage_cap) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= tiles) return;
    const SceneK& k = sc.scene(t);
    const uint32_t s0 = t * kTile;
    const uint32_t s1 = s0 + kTile < k.slot_end ? s0 + kTile : k.slot_end;
    write_region(sc, t, cell_of_pos(k, pos[s0]), cell_of_pos(k, pos[s1 - 1]), starts, out, stage_cap);
}

// A tile's region as uniform scalars (SGPRs).
struct RegionS {
    int32_t d0, d1, d2, d3, d4, d5, d6, d7, d8;
    uint32_t b1, b2, b3, b4, b5, b6, b7, b8;
    uint32_t total;
    bool staged;
    uint32_t scene, slot_end, active, refine, walls;
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
    s.scene = uni(R->scene);
    s.slot_end = uni(R->slot_end);
    s.active = uni(R->active);
    s.refine = uni(R->refine);
    s.walls = uni(R->walls);
    s.staged = uni(R->staged) != 0u;
    return s;
}

// Slot o's wall gradient (component `lane`, or whole): wall_grad for a tile
// with a particle near a wall, else the zero rows. A pointer select, not a
// branch, so the load is issued like the others.
__device__ __forceinline__ float wall_part(const SolveBufs& b, const RegionS& R, uint32_t o, uint32_t lane) {
    const float4* src = R.walls ? b.wall_grad : b.zeros;
    return lane_part(src, R.walls ? o : (o & (kZeroRows - 1u)), lane);
}
__device__ __forceinline__ float4 wall_grad_of(const SolveBufs& b, const RegionS& R, uint32_t o) {
    const float4* src = R.walls ? b.wall_grad : b.zeros;
    return src[R.walls ? o : (o & (kZeroRows - 1u))];
}

// The nine rows of a tile's region as lanes 0..8 of a wave, taken from the
// region's scalar registers: lane r holds row r's first LDS index and its
// delta (a slot j of row r sits at LDS index j + delta). Lanes past 8 are
// never read.
struct RowLanes {
    uint32_t base;
    int32_t delta;
};
__device__ __forceinline__ RowLanes row_lanes(const RegionS& R) {
    const uint32_t l = __lane_id();
    // Independent selects (each a compare and a conditional move), no
    // branches.
    const uint32_t bs[9] = {0u, R.b1, R.b2, R.b3, R.b4, R.b5, R.b6, R.b7, R.b8};
    const int32_t ds[9] = {R.d0, R.d1, R.d2, R.d3, R.d4, R.d5, R.d6, R.d7, R.d8};
    RowLanes T{0u, R.d0};
#pragma unroll
    for (uint32_t r = 1; r < kRows; ++r) {
        const bool here = l == r;
        T.base = here ? bs[r] : T.base;
        T.delta = here ? ds[r] : T.delta;
    }
    return T;
}
// The row holding flat LDS index i (uniform): the last row whose base <= i
// (empty rows share the next row's base, so that is the non-empty one).
__device__ __forceinline__ uint32_t row_of(const RowLanes& T, uint32_t i) {
    return (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(T.base <= i) & 0x1FEull);
}

// Copies the region's `total` slots into LDS: NT threads, ITERS slots each,
// all loads issued before the stores. `load(slot, it)` fetches into
// registers, `store(idx, it)` writes them to LDS. `early()` runs once, after
// the first chunk's loads are issued: the sweep's own loads placed there are
// in flight while the stores and the barrier complete.
//
// Each wave copies 64 consecutive LDS indices at a time. Its first and last
// index name the rows the 64 span (uniform), usually one or two: the wave
// takes the first row's delta and, per row boundary inside the span, one
// compare and select a lane, instead of each thread searching the nine rows.
//
// Every load and store is issued unconditionally, so the compiler knows how
// many loads are in flight and waits for the staging loads alone before the
// stores. A thread past the region re-reads the region's last slot and
// stores it at its own index, which the tile leaves unused (distinct
// addresses, so no bank serialisation); an index past the array's `slots`
// stores into its last slot.
template <uint32_t NT, uint32_t ITERS, class Load, class Store, class Early>
__device__ __forceinline__ void stage_rows(const RegionS& R, uint32_t slots, Load&& load, Store&& store,
                                           Early&& early) {
    if (R.total == 0) {
        early();
        return;
    }
    const RowLanes T = row_lanes(R);
    const uint32_t lane = __lane_id(), wave0 = uni(threadIdx.x & ~63u), last = R.total - 1u;
    auto chunk_loads = [&](uint32_t chunk) {
#pragma unroll
        for (uint32_t it = 0; it < ITERS; ++it) {
            const uint32_t i0 = chunk + it * NT + wave0;  // the wave's first index (uniform)
            const uint32_t f = i0 < last ? i0 : last, l = i0 + 63u < last ? i0 + 63u : last;
            const uint32_t r0 = row_of(T, f), r1 = row_of(T, l);
            const uint32_t ic = i0 + lane < last ? i0 + lane : last;
            int32_t d = __builtin_amdgcn_readlane(T.delta, (int)r0);
            for (uint32_t r = r0 + 1u; r <= r1; ++r) {
                const uint32_t br = (uint32_t)__builtin_amdgcn_readlane((int)T.base, (int)r);
                const int32_t dr = __builtin_amdgcn_readlane(T.delta, (int)r);
                d = ic >= br ? dr : d;
            }
            load((uint32_t)((int32_t)ic - d), it);
        }
    };
    auto chunk_stores = [&](uint32_t chunk) {
#pragma unroll
        for (uint32_t it = 0; it < ITERS; ++it) {
            const uint32_t idx = chunk + it * NT + threadIdx.x;
            store(idx < slots ? idx : slots - 1u, it);
        }
    };
    chunk_loads(0);
    early();
    chunk_stores(0);
    for (uint32_t chunk = NT * ITERS; chunk < R.total; chunk += NT * ITERS) {
        chunk_loads(chunk);
        chunk_stores(chunk);
    }
}
struct NoEarly {
    __device__ void operator()() const {}
};


__device__ __forceinline__ f3 listed_grad(f3 pi, f3 pj, float g) { return scl(g, sub(pi, pj)); }


struct LaneId {
    uint32_t tile, p, lane, gid;
    bool live;
};

// The particle and lane of this thread; `live` when the slot holds one of
// the scene's particles (slots past them pad the scene to whole tiles).
__device__ __forceinline__ LaneId lane_id(uint32_t tile, uint32_t slot_end) {
    LaneId id;
    id.tile = tile;
    id.gid = id.tile * kSweepThreads + threadIdx.x;  // = p * kLanes + lane
    id.p = id.tile * kTile + threadIdx.x / kLanes;
    id.lane = threadIdx.x % kLanes;
    id.live = id.p < slot_end;
    return id;
}


// Lane `lane` of particle p walks its share of the list in order (entries
// lane, lane + 4, ...), calling f(neighbour, dW/r): the neighbour is an LDS
// index in a staged tile, a slot otherwise. The entries of the fixed rows
// come with the head; the rest, from the wave's tail block, kPrefetch quads
// at a time.
//
// Every list load takes a base that is uniform over the wave (a fixed row's
// start, or the wave's tail block advanced by the pass through it) plus a
// 32-bit byte offset the lane keeps, so it needs no address arithmetic of its
// own. A tail load may run up to kPrefetch - 1 quads past the lane's last
// entry: into later entries of the wave's block, or past the last block into
// the list's padding (kListPad), and never used.
template <bool STAGED>
using ListIdx = typename std::conditional<STAGED, uint16_t, uint32_t>::type;
template <bool STAGED>
__device__ __forceinline__ const char* list_ids(const SolveBufs& b, size_t e) {
    return STAGED ? reinterpret_cast<const char*>(b.nbr16 + e) : reinterpret_cast<const char*>(b.nbr32 + e);
}
template <bool STAGED>
__device__ __forceinline__ uint32_t list_id_at(const char* base, uint32_t byte_off) {
    return *reinterpret_cast<const ListIdx<STAGED>*>(base + byte_off);
}
__device__ __forceinline__ float list_g_at(const char* base, uint32_t byte_off) {
    return *reinterpret_cast<const float*>(base + byte_off);
}
// The slot whose own fields a thread reads: its particle, or for a lane past
// the scene's end the tile's first slot (a tile that runs holds one), so the
// sweeps' own loads need no branch.
__device__ __forceinline__ uint32_t own_slot(const LaneId& id) { return id.live ? id.p : id.tile * kTile; }

// Where a particle's tail starts in its wave's block: the summed tail lengths
// of the wave's earlier particles. Every lane of the wave takes part, the
// four lanes of a particle with its count, a lane past the scene's end with 0.
__device__ __forceinline__ uint32_t tail_offset(uint32_t nn) {
    const uint32_t len = list_tail_len(nn);
    return wave_inclusive_sum((__lane_id() & (kLanes - 1)) == 0 ? len : 0u) - len;
}
// The first entry of the wave's tail block (uniform).
__device__ __forceinline__ uint32_t wave_tail_block(const SolveBufs& b, const LaneId& id) {
    return uni(list_tail(b.stride, b.cap, id.p & ~(kListWave - 1u), 0u));
}

// The particle's entry count and the lane's entries in the kHead fixed rows.
// A sweep loads them before it stages its tile, so their latency overlaps
// the staging. The loads are unconditional: rows past the count, and those
// of a lane past the scene's end, lie inside the list and are never used.
constexpr uint32_t kHead = kListFixed;
struct ListHead {
    uint32_t nn;     // the particle's entries (its lane's share is taken later)
    uint32_t wbase;  // the wave's tail block (uniform)
    uint32_t lane_off;  // the lane's first entry past the head, from wbase
    uint32_t j[kHead];
    float g[kHead];
};
template <bool STAGED>
__device__ __forceinline__ ListHead list_head(const SolveBufs& b, const LaneId& id) {
    ListHead h;
    h.nn = at32(b.nbr_n, own_slot(id));
    h.wbase = 0;
    h.lane_off = 0;
    const uint32_t last = b.cap / kLanes - 1u;
    const uint32_t vj = id.gid * (uint32_t)sizeof(ListIdx<STAGED>), vg = id.gid * 4u;
#pragma unroll
    for (uint32_t u = 0; u < kHead; ++u) {
        const size_t row = (size_t)(u < last ? u : last) * b.stride * kLanes;
        h.j[u] = list_id_at<STAGED>(list_ids<STAGED>(b, row), vj);
        h.g[u] = list_g_at(reinterpret_cast<const char*>(b.nbrg + row), vg);
    }
    return h;
}
// Sets the lane's tail place once the count has arrived; every lane of the
// wave calls it.
__device__ __forceinline__ void list_tail_at(const SolveBufs& b, const LaneId& id, ListHead& h) {
    h.wbase = wave_tail_block(b, id);
    h.lane_off = tail_offset(id.live ? h.nn : 0u) + id.lane;
}
// Calls f(j, g) over the lane's entries in order: the head, then the tail
// kPrefetch quads at a time. EARLY issues the first tail quads before the
// head is used, so their latency overlaps the head's pair math (where the
// registers allow it without costing occupancy).
template <bool STAGED, bool EARLY = false, class F>
__device__ __forceinline__ void for_list(const SolveBufs& b, const LaneId& id, const ListHead& h, F&& f) {
    const uint32_t cnt = id.live && h.nn > id.lane ? (h.nn - id.lane + kLanes - 1) / kLanes : 0u;
    uint32_t j[kPrefetch];
    float g[kPrefetch];
    const uint32_t vj = h.lane_off * (uint32_t)sizeof(ListIdx<STAGED>), vg = h.lane_off * 4u;
    auto load = [&](uint32_t t) {
        // t is uniform: the quad's entries start kLanes (t - kHead) past the
        // block's base.
        const size_t e = (size_t)h.wbase + (size_t)(t - kHead) * kLanes;
        const char* jb = list_ids<STAGED>(b, e);
        const char* gb = reinterpret_cast<const char*>(b.nbrg + e);
#pragma unroll
        for (uint32_t u = 0; u < kPrefetch; ++u) {
            j[u] = list_id_at<STAGED>(jb + u * kLanes * sizeof(ListIdx<STAGED>), vj);
            g[u] = list_g_at(gb + u * kLanes * sizeof(float), vg);
        }
    };
    if (EARLY && cnt > kHead) load(kHead);
#pragma unroll
    for (uint32_t u = 0; u < kHead; ++u)
        if (u < cnt) f(h.j[u], h.g[u]);
    for (uint32_t t = kHead; t < cnt; t += kPrefetch) {
        if (!EARLY || t > kHead) load(t);
#pragma unroll
        for (uint32_t u = 0; u < kPrefetch; ++u)
            if (t + u < cnt) f(j[u], g[u]);
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
    const float r = sqrt_rn(r2);
    const float q = r * k.inv_h;
    a.w_sum += kernel_q(k, q);
    const float wn = wmax(1.0f - q, 0.0f);
    a.rn += wn * wn * wn;
    const float g = r < 1e-9f ? 0.0f : div_rn(grad_dw(k, q), r);
    const f3 gw = r < 1e-9f ? mk(0.0f, 0.0f, 0.0f) : scl(g, x);
    a.gsum = add(a.gsum, gw);
    a.gsq += dot3(gw, gw);
    a.drho += dot3(sub(vi, vj), gw);
    return g;
}

// Distances from the particle to the faces of its own cell: a stencil row or
// end cell farther than the support holds no neighbour at a finite
// position, so the walk skips it (the accepted neighbours and their order
// are those of the full walk: a skipped candidate would fail the distance
// test). A non-finite position has an unordered distance to every
// candidate, which the shader keeps (sim_solve.wgsl:450), so a particle
// with one, and every particle of a scene that holds one (`guard`), walks
// its whole stencil.
struct Reach {
    float dl[3], dh[3];
    float lim;
};

__device__ __forceinline__ Reach make_reach(const SceneK& k, f3 p, uint32_t bx, uint32_t by, uint32_t bz,
                                            bool guard) {
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
    // A margin over the rounding of the faces.
    const bool finite = __builtin_isfinite(p.x) && __builtin_isfinite(p.y) && __builtin_isfinite(p.z);
    st.lim = finite && !guard ? k.support_sq * 1.001f : __builtin_huge_valf();
    return st;
}

// A 16-bit LDS word, typed so that its loads and stores stay LDS instructions.
typedef __attribute__((address_space(3))) uint16_t LdsU16;

// LDS layout of density_div and den_kappa: per slot the position and the
// velocity, 24 bytes. density_div adds its walk list after them.
struct PvLds {
    float4* pv;  // x, y, z, vx
    float2* vw;  // vy, vz
    LdsU16* list;    // density_div: kWalkList entries a particle of the tile
};
constexpr size_t kPvEntry = sizeof(float4) + sizeof(float2);  // 24

// A staged tile's walk writes each particle's first kWalkList entries (LDS
// indices) to LDS, at the particle's base plus the entry's place, and the
// pair sums move them to the list in its quad layout; entries past them go
// to the list directly. Developed flow at 1M particles holds at most 42.
constexpr uint32_t kWalkList = 48;
__host__ __device__ __forceinline__ uint32_t walk_list(const SolveBufs& b) {
    return b.cap < kWalkList ? b.cap : kWalkList;
}
__host__ __device__ __forceinline__ size_t density_lds(const SolveBufs& b) {
    return stage_slots(b) * kPvEntry + (size_t)kTile * walk_list(b) * sizeof(uint16_t);
}

template <bool STAGED>
__device__ __forceinline__ void load_pv(const PvLds& L, const float4* __restrict__ pos,
                                        const float4* __restrict__ vel, uint32_t j, f3& p, f3& v) {
    if (STAGED) {
        const float4 a = L.pv[j];
        const float2 b = L.vw[j];
        p = mk(a.x, a.y, a.z);
        v = mk(a.w, b.x, b.y);
    } else {
        p = xyz(pos[j]);
        v = xyz(vel[j]);
    }
}

template <uint32_t NT, uint32_t IT, class Early = NoEarly>
__device__ __forceinline__ void stage_pv(const RegionS& R, const float4* __restrict__ pos,
                                         const float4* __restrict__ vel, const PvLds& L, uint32_t slots,
                                         Early&& early = Early()) {
    float4 pp[IT], vv[IT];
    stage_rows<NT, IT>(
        R, slots,
        [&](uint32_t slot, uint32_t it) {
            pp[it] = at32(pos, slot);
            vv[it] = at32(vel, slot);
        },
        [&](uint32_t idx, uint32_t it) {
            L.pv[idx] = make_float4(pp[it].x, pp[it].y, pp[it].z, vv[it].x);
            L.vw[idx] = make_float2(vv[it].y, vv[it].z);
        },
        early);
}

// Phase 1 of density_div: the four lanes of particle i test consecutive
// candidates of each stencil row; a ballot orders the accepted ones, and each
// is appended (as its LDS index when staged, else its slot) at its place in
// walk order. Returns the neighbour count, past the cap included.
template <bool STAGED>
__device__ __forceinline__ uint32_t density_walk(const SceneK& k, const LaneId& id, f3 pi, uint32_t bx, uint32_t by,
                                                 uint32_t bz, const uint32_t* __restrict__ starts,
                                                 const float4* __restrict__ pos, const PvLds& L,
                                                 const RegionS& R, const SolveBufs& b, bool guard) {
    const Reach st = make_reach(k, pi, bx, by, bz, guard);
    const uint32_t shift = __lane_id() & ~(kLanes - 1);
    const uint32_t below = (1u << id.lane) - 1u;
    // Every row's candidate range first, so the 18 bound loads overlap.
    uint32_t k0s[kRows], k1s[kRows];
#pragma unroll
    for (uint32_t r = 0; r < kRows; ++r) {
        const int dz = (int)(r / 3) - 1, dy = (int)(r % 3) - 1;
        const int cz = (int)bz + dz, cy = (int)by + dy;
        const float ez = dz < 0 ? st.dl[2] : (dz > 0 ? st.dh[2] : 0.0f);
        const float ey = dy < 0 ? st.dl[1] : (dy > 0 ? st.dh[1] : 0.0f);
        const float e2 = ey * ey + ez * ez;
        const bool live_row = cz >= 0 && cz < (int)k.dims[2] && cy >= 0 && cy < (int)k.dims[1] && e2 <= st.lim;
        const uint32_t x0 = (bx > 0 && st.dl[0] * st.dl[0] + e2 <= st.lim) ? bx - 1 : bx;
        const uint32_t x1 = (bx + 1 < k.dims[0] && st.dh[0] * st.dh[0] + e2 <= st.lim) ? bx + 1 : bx;
        const uint32_t row = live_row ? k.cell_offset + ((uint32_t)cz * k.dims[1] + (uint32_t)cy) * k.dims[0] : 0u;
        k0s[r] = live_row ? starts[row + x0] : 0u;
        k1s[r] = live_row ? starts[row + x1 + 1] : 0u;
    }
    // List indices fit 32 bits (cap x stride < 2^32, checked at allocation).
    const uint32_t lbase = id.p * kLanes, lstep = b.stride * kLanes;
    const uint32_t T = walk_list(b);
    LdsU16* const mine = L.list + (threadIdx.x / kLanes) * T;
    // Staged, entries below T go to LDS; unstaged, those of the fixed rows go
    // to the list. The rest, up to the cap, wait in the spill area, slot-major
    // from the first tail entry: their place depends on the wave's counts.
    // (A cap below the fixed rows leaves only cap / 4 of them in the list.)
    const uint32_t fixed = kListFixed * kLanes, low = STAGED ? T : (fixed < b.cap ? fixed : b.cap);
    const uint32_t spill = id.p * list_tail_cap(b.cap) - fixed;
    auto append = [&](bool acc, uint32_t m, uint32_t j) {
        if (!acc) return;
        if (m < low) {
            if (STAGED)
                mine[m] = (uint16_t)j;
            else
                b.nbr32[(m / kLanes) * lstep + lbase + m % kLanes] = j;
        } else if (m < b.cap) {
            b.spill[spill + m] = j;
        }
    };
    uint32_t found = 0;
#pragma unroll
    for (uint32_t r = 0; r < kRows; ++r) {
        const uint32_t k0 = k0s[r], k1 = k1s[r];
        const int32_t delta = !STAGED ? 0
                              : r == 0 ? R.d0
                              : r == 1 ? R.d1
                              : r == 2 ? R.d2
                              : r == 3 ? R.d3
                              : r == 4 ? R.d4
                              : r == 5 ? R.d5
                              : r == 6 ? R.d6
                              : r == 7 ? R.d7
                                       : R.d8;
        // Two quads of candidates a trip (kq + lane, kq + 4 + lane), tested
        // branch-free, a lane past the row's end masked out of the ballot.
        // Staged, it reads at most 7 slots past the row's LDS range, inside
        // the array's spare slots; unstaged, it re-reads the row's last slot.
        // Only the centre row holds the particle itself.
        const uint32_t last = k1 - 1u;
        for (uint32_t kq = k0; kq < k1; kq += 2 * kLanes) {
            const uint32_t ka = kq + id.lane, kb = ka + kLanes;
            const bool ina = ka < k1, inb = kb < k1;
            const uint32_t ja = STAGED ? (uint32_t)((int32_t)ka + delta) : (ina ? ka : last);
            const uint32_t jb = STAGED ? (uint32_t)((int32_t)kb + delta) : (inb ? kb : last);
            f3 pa, pb;
            if (STAGED) {
                const float4 a = L.pv[ja], c = L.pv[jb];
                pa = mk(a.x, a.y, a.z);
                pb = mk(c.x, c.y, c.z);
            } else {
                pa = xyz(pos[ja]);
                pb = xyz(pos[jb]);
            }
            const f3 xa = sub(pi, pa), xb = sub(pi, pb);
            // The shader skips a candidate when r2 >= support_sq
            // (sim_solve.wgsl:450), so an unordered (NaN) distance is kept.
            const bool near_a = !(dot3(xa, xa) >= k.support_sq), near_b = !(dot3(xb, xb) >= k.support_sq);
            bool acc_a = ina & near_a;
            bool acc_b = inb & near_b;
            // Each ballot of a single compare is that compare's lane mask, so
            // the masks combine in scalar registers.
            uint64_t ba = __builtin_amdgcn_ballot_w64(ina) & __builtin_amdgcn_ballot_w64(near_a);
            uint64_t bb = __builtin_amdgcn_ballot_w64(inb) & __builtin_amdgcn_ballot_w64(near_b);
            if (r == 4) {
                acc_a &= ka != id.p;
                acc_b &= kb != id.p;
                ba &= __builtin_amdgcn_ballot_w64(ka != id.p);
                bb &= __builtin_amdgcn_ballot_w64(kb != id.p);
            }
            const uint32_t qa = (uint32_t)(ba >> shift) & 0xFu;
            const uint32_t qb = (uint32_t)(bb >> shift) & 0xFu;
            const uint32_t fb = found + __popc(qa);
            append(acc_a, found + __popc(qa & below), ja);
            append(acc_b, fb + __popc(qb & below), jb);
            found = fb + __popc(qb);
        }
    }
    return found;
}

// Phase 2: lane l's share of the pair sums over the list (entries l, l + 4,
// ...), storing each pair's neighbour and dW/r at its place in the list
// (`tail`: the lane's first entry past the fixed rows); past the cap, the
// neighbours the list dropped are found again by a global walk and summed by
// the same rule.
template <bool STAGED>
__device__ __forceinline__ void density_sums(const SceneK& k, const LaneId& id, f3 pi, f3 vi, uint32_t found,
                                             uint32_t tail, uint32_t bx, uint32_t by, uint32_t bz,
                                             const uint32_t* __restrict__ starts, const float4* __restrict__ pos,
                                             const float4* __restrict__ vel, const PvLds& L, const SolveBufs& b,
                                             DensitySums& a) {
    const uint32_t nn = found < b.cap ? found : b.cap;
    const uint32_t cnt = nn > id.lane ? (nn - id.lane + kLanes - 1) / kLanes : 0u;
    const uint32_t step = b.stride * kLanes, last = b.cap / kLanes - 1u;
    auto at = [&](uint32_t t) { return t < kListFixed ? t * step + id.gid : tail + (t - kListFixed) * kLanes; };
    // The spill index of the lane's entry 4t + lane is spilled + 4t.
    const uint32_t spilled = id.p * list_tail_cap(b.cap) - kListFixed * kLanes + id.lane;
    auto pair = [&](uint32_t t, uint32_t j) {
        f3 pj, vj;
        load_pv<STAGED>(L, pos, vel, j, pj, vj);
        b.nbrg[at(t)] = density_pair(k, pi, vi, pj, vj, a);
    };
    if (STAGED) {
        // The lane's first T/4 entries come from the walk's LDS list, any
        // further ones from the spill area. Two loops, so the LDS one waits
        // on no global load (or, through it, on the stores).
        const uint32_t T = walk_list(b);
        const LdsU16* const mine = L.list + (threadIdx.x / kLanes) * T + id.lane;
        const uint32_t in_lds = cnt < T / kLanes ? cnt : T / kLanes;
        for (uint32_t t = 0; t < in_lds; ++t) {
            const uint32_t j = mine[t * kLanes];
            b.nbr16[at(t)] = (uint16_t)j;
            pair(t, j);
        }
        for (uint32_t t = in_lds; t < cnt; ++t) {
            const uint32_t j = b.spill[spilled + t * kLanes];
            b.nbr16[at(t)] = (uint16_t)j;
            pair(t, j);
        }
    } else {
        // The fixed rows hold the lane's first kHead entries already; they
        // load together (unconditionally: rows past the count lie inside the
        // list and are not used). Further entries come from the spill area.
        uint32_t jh[kHead];
#pragma unroll
        for (uint32_t u = 0; u < kHead; ++u) jh[u] = b.nbr32[(u < last ? u : last) * step + id.gid];
#pragma unroll
        for (uint32_t u = 0; u < kHead; ++u)
            if (u < cnt) pair(u, jh[u]);
        for (uint32_t t = kHead; t < cnt; ++t) {
            const uint32_t j = b.spill[spilled + t * kLanes];
            b.nbr32[at(t)] = j;
            pair(t, j);
        }
    }
    if (found > b.cap) {
        uint32_t seen = 0;
        for (uint32_t rr = 0; rr < kRows; ++rr) {
            const int dz = (int)(rr / 3) - 1, dy = (int)(rr % 3) - 1;
            const int cz = (int)bz + dz, cy = (int)by + dy;
            if (cz < 0 || cz >= (int)k.dims[2] || cy < 0 || cy >= (int)k.dims[1]) continue;
            const uint32_t x0 = bx > 0 ? bx - 1 : bx;
            const uint32_t x1 = bx + 1 < k.dims[0] ? bx + 1 : bx;
            const uint32_t row = k.cell_offset + ((uint32_t)cz * k.dims[1] + (uint32_t)cy) * k.dims[0];
            const uint32_t k0 = starts[row + x0], k1 = starts[row + x1 + 1];
            for (uint32_t kk = k0; kk < k1; ++kk) {
                const f3 pj = xyz(pos[kk]);
                const f3 x = sub(pi, pj);
                if (!(dot3(x, x) >= k.support_sq) && kk != id.p) {
                    if (seen >= b.cap && seen % kLanes == id.lane) density_pair(k, pi, vi, pj, xyz(vel[kk]), a);
                    seen++;
                }
            }
        }
    }
}

// The workgroup's sum of the particles' fixed-point errors: a wave sum, the waves' sums through LDS, one atomic. Integer sums, so the total
// is the same in any order.
__device__ __forceinline__ void block_residual(unsigned long long v, unsigned long long* out) {
    __shared__ unsigned long long s_wave[kSweepThreads / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += (unsigned long long)__shfl_xor((long long)v, o);
    if (__lane_id() == 0) s_wave[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
#pragma unroll
        for (uint32_t w = 0; w < kSweepThreads / 64; ++w) t += s_wave[w];
        if (t) atomicAdd(out, t);
    }
}

// A particle's error e as a 2^-32 fixed-point integer (tolerance mode),
// saturated at 2^40 (e = 256): NaN and non-positive errors count zero.
__device__ __forceinline__ unsigned long long fixed_err(float e) {
    e = e * 4294967296.0f;
    return e > 0.0f ? (unsigned long long)(e < 1099511627776.0f ? e : 1099511627776.0f) : 0ull;
}

// Lane 0's end of density_div for particle i: the wall terms, ρ, α and the
// divergence κ/ρ (with TOL, its fixed-point divergence error too). Returns
// whether the particle's wall gradient is nonzero.
template <bool TOL>
__device__ __forceinline__ bool density_finish(const SceneK& k, const StepK& s, uint32_t* counters, const LaneId& id,
                                               f3 pi, f3 vi, uint32_t found, float w_sum, float rn, float gsq_sum,
                                               float drho_sum, f3 gsum_sum, const SolveBufs& b,
                                               unsigned long long& err) {
    const uint32_t i = id.p;
    const uint32_t nn = found < b.cap ? found : b.cap;
    b.nbr_n[i] = nn;
    if (found > b.cap) atomicAdd(&counters[1], found - b.cap);
    float rho = k.mass * w_sum;
    const float gsq = k.mass * k.mass * gsq_sum;
    float drho = k.mass * drho_sum;
    f3 gsum = scl(k.mass, gsum_sum);
    // The wall terms, which a wave with no particle near a wall skips (they
    // would all be +0).
    float fill = 0.0f;
    f3 wg = mk(0.0f, 0.0f, 0.0f);
    if (near_walls(k, pi)) {
        const f3 tl = mk((pi.x - k.wall_lo[0]) * k.inv_h, (pi.y - k.wall_lo[1]) * k.inv_h,
                         (pi.z - k.wall_lo[2]) * k.inv_h);
        const f3 th = mk((k.wall_hi[0] - pi.x) * k.inv_h, (k.wall_hi[1] - pi.y) * k.inv_h,
                         (k.wall_hi[2] - pi.z) * k.inv_h);
        fill += wall_density(tl.x) + wall_density(th.x);
        fill += wall_density(tl.y) + wall_density(th.y);
        fill += wall_density(tl.z) + wall_density(th.z);
        fill -= wedge_sum(tl, th);
        wg = wall_grad_sum(k, pi);
    }
    rho += k.rho0 * fill;
    b.wall_grad[i] = make_float4(wg.x, wg.y, wg.z, 0.0f);
    gsum = add(gsum, wg);
    b.density[i] = rho;
    b.rho_near[i] = rn;
    const float al = rho / wmax(dot3(gsum, gsum) + gsq, 1e-4f);
    b.alpha[i] = al;
    drho += dot3(vi, wg);
    b.kd[i] = wmax(drho, 0.0f) * al / s.dt_guard / rho;
    if (TOL) err = fixed_err(wmax(drho, 0.0f) * s.dt / k.rho0);
    return wg.x != 0.0f || wg.y != 0.0f || wg.z != 0.0f;
}

template <bool STAGED, bool TOL>
__device__ __forceinline__ void density_tile(const SceneK& k, const StepK& s, uint32_t* counters, const LaneId& id,
                                             const RegionS& R, const uint32_t* __restrict__ starts,
                                             const float4* __restrict__ pos, const float4* __restrict__ vel,
                                             const SolveBufs& b, const PvLds& L, unsigned long long& err) {
    if (STAGED) {
        stage_pv<kSweepThreads, 3>(R, pos, vel, L, stage_slots(b));
        __syncthreads();
    }
    f3 pi = mk(0.0f, 0.0f, 0.0f), vi = pi;
    uint32_t bx = 0, by = 0, bz = 0, found = 0;
    if (id.live) {
        pi = xyz(pos[id.p]);
        vi = xyz(vel[id.p]);
        bx = cell_coord(pi.x, k.box_min[0], k.cell, k.dims[0]);
        by = cell_coord(pi.y, k.box_min[1], k.cell, k.dims[1]);
        bz = cell_coord(pi.z, k.box_min[2], k.cell, k.dims[2]);
        found = density_walk<STAGED>(k, id, pi, bx, by, bz, starts, pos, L, R, b, counters[0] != 0u);
    }
    // Phase 1's list entries, written by any lane of the quad, are read back
    // by lane m % 4 below.
    __syncthreads();
    const uint32_t nn = id.live ? (found < b.cap ? found : b.cap) : 0u;
    const uint32_t tail = list_tail(b.stride, b.cap, id.p, tail_offset(nn)) + id.lane;
    DensitySums a{id.lane == 0 ? k.sigma : 0.0f, 0.0f, 0.0f, 0.0f, mk(0.0f, 0.0f, 0.0f)};
    if (id.live) density_sums<STAGED>(k, id, pi, vi, found, tail, bx, by, bz, starts, pos, vel, L, b, a);
    const float w_sum = fold4(a.w_sum), rn = fold4(a.rn), gsq_sum = fold4(a.gsq), drho_sum = fold4(a.drho);
    const f3 gsum_sum = fold4(a.gsum);
    if (!id.live || id.lane != 0) return;
    // A particle near a wall marks its tile (k_tile_regions cleared the mark
    // this substep; any number of lanes may store the same 1).
    if (density_finish<TOL>(k, s, counters, id, pi, vi, found, w_sum, rn, gsq_sum, drho_sum, gsum_sum, b, err))
        b.regions[id.tile].walls = 1u;
}


// At most 80 VGPRs (six waves a SIMD: three 512-thread workgroups a CU); the
// scene-table form would take a few more and drop to two, which costs it
// 16% against a 16-byte spill. TOL (tolerance mode) adds the divergence
// error of the solve's first pass.
template <class Sc, bool TOL = false>
__global__ __launch_bounds__(kSweepThreads) __attribute__((amdgpu_waves_per_eu(6))) void k_density_div(Sc sc, const uint32_t* __restrict__ starts,
                                                               const float4* __restrict__ pos,
                                                               const float4* __restrict__ vel, SolveBufs b) {
    extern __shared__ float4 smem[];
    const uint32_t tile = tile_index(sc);
    const RegionS R = load_region(b.regions, tile);
    if (!sc.runs_at(R.scene, R.active, R.refine, kEveryPass)) return;
    const SceneK& k = sc.scene_at(R.scene);
    const StepK& s = sc.step_at(R.scene);
    const LaneId id = lane_id(tile, sc.slot_end_at(R.slot_end));
    uint32_t* counters = b.counters + 2 * sc.index_at(R.scene);
    const PvLds L{smem, reinterpret_cast<float2*>(smem + stage_slots(b)),
                  (LdsU16*)(reinterpret_cast<char*>(smem) + stage_slots(b) * kPvEntry)};
    unsigned long long err = 0;
    if (R.staged)
        density_tile<true, TOL>(k, s, counters, id, R, starts, pos, vel, b, L, err);
    else
        density_tile<false, TOL>(k, s, counters, id, R, starts, pos, vel, b, L, err);
    if (TOL) block_residual(err, resid_word(b.div_residual, sc.index_at(R.scene), tile, b.resid_slots));
}

// ---- list sweeps -------------------------------------------------------------

// Stages position and kd (the gathers of div_apply and den_apply) as one
// float4 per slot.
template <class Early>
__device__ __forceinline__ void stage_pos_kd(const RegionS& R, const float4* __restrict__ pos,
                                             const float* __restrict__ kd, float4* lds, uint32_t slots,
                                             Early&& early) {
    constexpr uint32_t IT = 3;
    float4 pp[IT];
    float kk[IT];
    stage_rows<kSweepThreads, IT>(
        R, slots,
        [&](uint32_t slot, uint32_t it) {
            pp[it] = at32(pos, slot);
            kk[it] = at32(kd, slot);
        },
        [&](uint32_t idx, uint32_t it) { lds[idx] = make_float4(pp[it].x, pp[it].y, pp[it].z, kk[it]); }, early);
}

// Lane `lane` of particle p: the partial sum of (kd_i + kd_j) grad W_ij over
// its share of the list (kappa_partial, sim_solve.wgsl:596-606).
template <bool STAGED>
__device__ __forceinline__ f3 kappa_partial(const SceneK& k, const SolveBufs& b, const LaneId& id, const ListHead& h,
                                            f3 pi, float kdi, const float4* __restrict__ pos, const float4* lds) {
    f3 dv = mk(0.0f, 0.0f, 0.0f);
    for_list<STAGED, true>(b, id, h, [&](uint32_t j, float g) {
        float4 q;
        if (STAGED) {
            q = lds[j];
        } else {
            const float4 pj = pos[j];
            q = make_float4(pj.x, pj.y, pj.z, b.kd[j]);
        }
        dv = add(dv, scl(kdi + q.w, listed_grad(pi, xyz(q), g)));
    });
    return dv;
}

// `own(slot)` loads the caller's own fields with the list head.
template <bool STAGED, class Own>
__device__ __forceinline__ f3 kappa_sweep(const SceneK& k, const LaneId& id, const RegionS& R, const float4* __restrict__ pos,
                                          const SolveBufs& b, float4* smem, f3& pi, float& kdi, Own&& own) {
    ListHead h;
    float4 p4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float kd0 = 0.0f;
    auto early = [&] {
        const uint32_t o = own_slot(id);
        h = list_head<STAGED>(b, id);
        p4 = pos[o];
        kd0 = b.kd[o];
        own(o);
    };
    if (STAGED) {
        stage_pos_kd(R, pos, b.kd, smem, stage_slots(b), early);
        __syncthreads();
    } else {
        early();
    }
    list_tail_at(b, id, h);
    f3 dv = mk(0.0f, 0.0f, 0.0f);
    if (id.live) {
        pi = xyz(p4);
        kdi = kd0;
        dv = kappa_partial<STAGED>(k, b, id, h, pi, kdi, pos, smem);
    }
    return fold4(dv);
}

// div_apply / den_apply: v -= dt (m sum (kd_i + kd_j) gradW + kd_i wall_grad).
// Reads only kd of the neighbours, so the velocity update is in place
// (vel_out = vel_in) unless the caller keeps both (the taped forward).
template <class Sc>
__device__ __forceinline__ void kappa_apply_on(const Sc& sc, uint32_t tile, uint32_t pass,
                                               const float4* __restrict__ pos, const float4* vel_in,
                                               float4* vel_out, const SolveBufs& b, float4* smem) {
    const RegionS R = load_region(b.regions, tile);
    if (!sc.runs_at(R.scene, R.active, R.refine, pass)) return;
    const SceneK& k = sc.scene_at(R.scene);
    const StepK& s = sc.step_at(R.scene);
    const LaneId id = lane_id(tile, sc.slot_end_at(R.slot_end));
    f3 pi = mk(0.0f, 0.0f, 0.0f);
    float kdi = 0.0f;
    float wv = 0.0f, gv = 0.0f;  // component `lane` of vel_in and wall_grad
    auto own = [&](uint32_t o) {
        wv = lane_part(vel_in, o, id.lane);
        gv = wall_part(b, R, o, id.lane);
    };
    const f3 total = R.staged ? kappa_sweep<true>(k, id, R, pos, b, smem, pi, kdi, own)
                              : kappa_sweep<false>(k, id, R, pos, b, smem, pi, kdi, own);
    const float4 w = quad_float4(wv), wg = quad_float4(gv);
    if (id.live && id.lane == 0) {
        const f3 full = add(scl(k.mass, total), scl(kdi, xyz(wg)));
        vel_out[id.p] = make_float4(w.x - s.dt * full.x, w.y - s.dt * full.y, w.z - s.dt * full.z, w.w);
    }
}
template <class Sc>
__global__ __launch_bounds__(kSweepThreads) void k_kappa_apply(Sc sc, uint32_t pass, const float4* __restrict__ pos,
                                                               const float4* vel_in, float4* vel_out, SolveBufs b) {
    extern __shared__ float4 smem[];
    kappa_apply_on<Sc>(sc, tile_index(sc), pass, pos, vel_in, vel_out, b, smem);
}

// The forces sweep's gathers: position, velocity, density, near-density,
// temperature, m / rho and the warm start's kd = 0.5 pp / rho^2, 44 bytes a
// slot as float4 (x, y, z, vx), float4 (vy, vz, rho, rho_near), float2
// (temperature, m / rho) and float (kd). m / rho and kd are the neighbour's
// own quotients, so each staged slot divides once instead of each pair.
struct ForcesLds {
    float4* a;
    float4* c;
    float2* t;
    float* w;
};

struct ForceSums {
    f3 visc, blend, nearf, tension, warm;
    float heat;
};

// The particle's own fields the forces sweep ends with.
struct ForcesOwn {
    f3 pi;
    float4 v;     // vel_in, w kept
    float rho, prev_p, temp;
    float gv;     // component `lane` of wall_grad
};

// forces_eval's pair terms (sim_solve.wgsl:544-583) and, over the same list
// entries in the same order, the warm start's sum (kd_i + kd_j) grad W_ij of
// forces_den_apply (kappa_partial, sim_solve.wgsl:596-606, 642): the two
// kernels read the same list, and the warm start's kd is each particle's own
// 0.5 pp / rho^2, known before the sweep.
template <bool STAGED>
__device__ __forceinline__ void forces_tile(const SceneK& k, const LaneId& id, const RegionS& R,
                                            const float4* __restrict__ pos, const float4* __restrict__ vel,
                                            const float* __restrict__ temperature,
                                            const float* __restrict__ prev_pressure, const SolveBufs& b,
                                            const ForcesLds& L, ForceSums& F, ForcesOwn& own) {
    ListHead h;
    float4 p4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), v4 = p4;
    float w_i = 0.0f, temp_i = 0.0f, rho_i = 0.0f, prev_p = 0.0f, gv = 0.0f;
    auto early = [&] {
        const uint32_t o = own_slot(id);
        h = list_head<STAGED>(b, id);
        p4 = pos[o];
        v4 = vel[o];
        w_i = b.rho_near[o];
        temp_i = temperature[o];
        rho_i = b.density[o];
        prev_p = prev_pressure[o];
        gv = wall_part(b, R, o, id.lane);
    };
    if (STAGED) {
        constexpr uint32_t IT = 3;
        float4 pp[IT], vv[IT];
        float dd[IT], nr[IT], tt[IT], pq[IT];
        stage_rows<kSweepThreads, IT>(
            R, stage_slots(b),
            [&](uint32_t slot, uint32_t it) {
                pp[it] = at32(pos, slot);
                vv[it] = at32(vel, slot);
                dd[it] = at32(b.density, slot);
                nr[it] = at32(b.rho_near, slot);
                tt[it] = at32(temperature, slot);
                pq[it] = at32(prev_pressure, slot);
            },
            [&](uint32_t idx, uint32_t it) {
                L.a[idx] = make_float4(pp[it].x, pp[it].y, pp[it].z, vv[it].x);
                L.c[idx] = make_float4(vv[it].y, vv[it].z, dd[it], nr[it]);
                L.t[idx] = make_float2(tt[it], div_rn(k.mass, dd[it]));
                L.w[idx] = (0.5f * pq[it]) / (dd[it] * dd[it]);
            },
            early);
        __syncthreads();
    } else {
        early();
    }
    list_tail_at(b, id, h);
    own.pi = xyz(p4);
    own.v = v4;
    own.rho = rho_i;
    own.prev_p = prev_p;
    own.temp = temp_i;
    own.gv = gv;
    if (!id.live) return;
    const f3 vi = xyz(v4);
    const float rhoi = rho_i;
    const f3 pii = own.pi;
    const float kd_i = (0.5f * prev_p) / (rho_i * rho_i);
    for_list<STAGED, true>(b, id, h, [&](uint32_t j, float gs) {
        f3 pj, vj;
        float rho_j, near_j, temp_j, mr_j, kd_j;
        if (STAGED) {
            const float4 a = L.a[j], c = L.c[j];
            const float2 t = L.t[j];
            pj = mk(a.x, a.y, a.z);
            vj = mk(a.w, c.x, c.y);
            rho_j = c.z;
            near_j = c.w;
            temp_j = t.x;
            mr_j = t.y;
            kd_j = L.w[j];
        } else {
            pj = xyz(pos[j]);
            vj = xyz(vel[j]);
            rho_j = b.density[j];
            near_j = b.rho_near[j];
            temp_j = temperature[j];
            mr_j = div_rn(k.mass, rho_j);
            kd_j = (0.5f * prev_pressure[j]) / (rho_j * rho_j);
        }
        const f3 x = sub(pii, pj);
        const float r2 = dot3(x, x);
        const float rr = sqrt_rn(r2);
        const f3 g = scl(gs, x);
        F.warm = add(F.warm, scl(kd_i + kd_j, g));
        const float f = div_rn(dot3(x, g), r2 + k.eps_h2);
        const float pair = div_rn(k.mass, rhoi * rho_j) * f;
        const f3 dv = sub(vi, vj);
        const float wn = wmax(1.0f - rr * k.inv_h, 0.0f);
        // x / max(r, 1e-5) and x / max(r, 1e-9) (sim_solve.wgsl:559, 567-568)
        // as x times one reciprocal of r, shared when r >= 1e-5.
        const float inv_rg = div_rn(1.0f, wmax(rr, 1e-9f));
        const float inv_rm = rr >= 1e-5f ? inv_rg : 1.0f / 1e-5f;
        const float sn = (w_i + near_j) * wn * wn;
        F.nearf = add(F.nearf, mk(sn * (x.x * inv_rm), sn * (x.y * inv_rm), sn * (x.z * inv_rm)));
        F.visc = add(F.visc, scl(k.visc2 * pair, dv));
        F.blend = sub(F.blend, scl(mr_j * kernel_q(k, rr * k.inv_h), dv));
        F.heat += k.heat_cond * pair * (temp_i - temp_j);
        F.heat -= k.heat_visc * pair * dot3(dv, dv);
        const float corr = div_rn(2.0f * k.rho0, rhoi + rho_j);
        const float cm = corr * k.mass * cohesion(rr, k.coh_c, k.coh_k, k.coh_near);
        F.tension = sub(F.tension, mk(cm * x.x * inv_rg, cm * x.y * inv_rg, cm * x.z * inv_rg));
    });
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

// forces_eval + forces_den_apply (sim_solve.wgsl:524-592, 630-663) in one
// sweep: the accelerations, the XSPH blend and the warm start's kd and
// pressure, then at once the particle's forced velocity less the warm-start
// correction. It reads the neighbours' velocities and temperatures, so the
// new velocity and temperature go to arrays of their own (vel_out,
// temp_out), not in place.
template <class Sc>
__global__ __launch_bounds__(kSweepThreads) void k_forces(Sc sc, const float4* __restrict__ pos,
                                                          const float4* __restrict__ vel,
                                                          const float* __restrict__ temperature,
                                                          const float* __restrict__ prev_pressure,
                                                          float4* __restrict__ vel_out,
                                                          float* __restrict__ temp_out, SolveBufs b) {
    extern __shared__ float4 smem[];
    const uint32_t tile = tile_index(sc);
    const RegionS R = load_region(b.regions, tile);
    if (!sc.runs_at(R.scene, R.active, R.refine, kEveryPass)) return;
    const SceneK& k = sc.scene_at(R.scene);
    const StepK& s = sc.step_at(R.scene);
    const LaneId id = lane_id(tile, sc.slot_end_at(R.slot_end));
    // The scene's first tile clears the grid build's non-finite flag: this
    // substep's walk has read it, and the resting set integrate writes is the
    // next build's to judge.
    if (threadIdx.x == 0 && tile * kTile == k.slot_begin) b.counters[2 * sc.index_at(R.scene)] = 0u;
    const uint32_t slots = stage_slots(b);
    float2* const t2 = reinterpret_cast<float2*>(smem + 2 * slots);
    const ForcesLds L{smem, smem + slots, t2, reinterpret_cast<float*>(t2 + slots)};
    const f3 z = mk(0.0f, 0.0f, 0.0f);
    ForceSums F{z, z, z, z, z, 0.0f};
    ForcesOwn own{};
    if (R.staged)
        forces_tile<true>(k, id, R, pos, vel, temperature, prev_pressure, b, L, F, own);
    else
        forces_tile<false>(k, id, R, pos, vel, temperature, prev_pressure, b, L, F, own);
    const f3 visc = fold4(F.visc), blend = fold4(F.blend), nearf = fold4(F.nearf), tension = fold4(F.tension);
    const f3 warm = fold4(F.warm);
    const float heat = fold4(F.heat);
    const float4 wg = quad_float4(own.gv);
    if (id.live && id.lane == 0) {
        const uint32_t i = id.p;
        const f3 pi = own.pi;
        const f3 adh = scl(k.beta * k.rho0, near_walls(k, pi) ? wall_adh_sum(k, pi) : mk(0.0f, 0.0f, 0.0f));
        const f3 acc = add(add(add(visc, scl(k.k_near, nearf)), scl(k.gamma, tension)), adh);
        b.accel[i] = make_float4(acc.x, acc.y, acc.z, heat);
        b.xsph[i] = make_float4(blend.x, blend.y, blend.z, 0.0f);
        const float half = 0.5f * own.prev_p;
        const float kdi = half / (own.rho * own.rho);
        b.kd[i] = kdi;
        b.pressure[i] = half;
        const f3 v = xyz(own.v);
        const f3 om = mk(s.omega[0], s.omega[1], s.omega[2]);
        const f3 dom = mk(s.domega[0], s.domega[1], s.domega[2]);
        const f3 c1 = cross3(dom, pi);
        const f3 c2 = cross3(om, cross3(om, pi));
        const f3 c3 = scl(2.0f, cross3(om, v));
        const f3 fict = sub(sub(mk(-c1.x, -c1.y, -c1.z), c2), c3);
        const f3 fsum = add(add(mk(s.force[0], s.force[1], s.force[2]), fict), acc);
        const f3 forced = fingers(s, pi, add(add(v, scl(s.dt, fsum)), scl(s.c_xsph, blend)));
        temp_out[i] = own.temp + s.dt * heat;
        const f3 full = add(scl(k.mass, warm), scl(kdi, xyz(wg)));
        vel_out[i] = make_float4(forced.x - s.dt * full.x, forced.y - s.dt * full.y, forced.z - s.dt * full.z,
                                 own.v.w);
    }
}

// The particle's own fields den_kappa ends with, loaded with the list head.
struct DenOwn {
    float rho, alpha, pressure;
    float4 wg;
};

template <bool STAGED>
__device__ __forceinline__ float den_kappa_sweep(const SceneK& k, const LaneId& id, const RegionS& R, const float4* __restrict__ pos,
                                                 const float4* __restrict__ vel, const SolveBufs& b, const PvLds& L,
                                                 f3& vi, DenOwn& own) {
    ListHead h;
    float4 p4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), v4 = p4;
    auto early = [&] {
        const uint32_t o = own_slot(id);
        h = list_head<STAGED>(b, id);
        p4 = pos[o];
        v4 = vel[o];
        own.rho = b.density[o];
        own.alpha = b.alpha[o];
        own.pressure = b.pressure[o];
        own.wg = wall_grad_of(b, R, o);
    };
    if (STAGED) {
        stage_pv<kSweepThreads, 3>(R, pos, vel, L, stage_slots(b), early);
        __syncthreads();
    } else {
        early();
    }
    list_tail_at(b, id, h);
    float drho = 0.0f;
    if (id.live) {
        const f3 pi = xyz(p4);
        vi = xyz(v4);
        const f3 vii = vi;
        for_list<STAGED, true>(b, id, h, [&](uint32_t j, float g) {
            f3 pj, vj;
            load_pv<STAGED>(L, pos, vel, j, pj, vj);
            drho += dot3(sub(vii, vj), listed_grad(pi, pj, g));
        });
    }
    return fold4(drho);
}

// den_kappa (sim_solve.wgsl:674-698), and with DIV the divergence kappa of
// tolerance mode's further divergence passes: density_div's kappa formula
// (sim_solve.wgsl:508-516) from the current velocities, no pressure.
template <class Sc, bool DIV>
__device__ __forceinline__ void den_kappa_on(const Sc& sc, uint32_t tile, uint32_t pass,
                                             const float4* __restrict__ pos, const float4* __restrict__ vel,
                                             const SolveBufs& b, float4* smem) {
    const RegionS R = load_region(b.regions, tile);
    if (!sc.runs_at(R.scene, R.active, R.refine, pass)) return;
    const SceneK& k = sc.scene_at(R.scene);
    const StepK& s = sc.step_at(R.scene);
    const LaneId id = lane_id(tile, sc.slot_end_at(R.slot_end));
    const PvLds L{smem, reinterpret_cast<float2*>(smem + stage_slots(b)), nullptr};
    f3 vi = mk(0.0f, 0.0f, 0.0f);
    DenOwn own;
    const float total = R.staged ? den_kappa_sweep<true>(k, id, R, pos, vel, b, L, vi, own)
                                 : den_kappa_sweep<false>(k, id, R, pos, vel, b, L, vi, own);
    unsigned long long err = 0;
    unsigned long long* const sum = DIV ? b.div_residual : b.residual;
    if (id.live && id.lane == 0) {
        const uint32_t i = id.p;
        const float rho = own.rho;
        const float d = k.mass * total + dot3(vi, xyz(own.wg));
        if (DIV) {
            b.kd[i] = wmax(d, 0.0f) * own.alpha / s.dt_guard / rho;
            err = fixed_err(wmax(d, 0.0f) * s.dt / k.rho0);
        } else {
            const float rho_star = rho + s.dt * d;
            const float excess = wmax(rho_star - k.rho0, 0.0f);
            const float kap = excess * own.alpha / (s.dt * s.dt);
            b.kd[i] = kap / rho;
            b.pressure[i] = own.pressure + kap * rho;
            if (sum) err = fixed_err(excess / k.rho0);
        }
    }
    if (sum) block_residual(err, resid_word(sum, sc.index_at(R.scene), tile, b.resid_slots));
}
template <class Sc, bool DIV = false>
__global__ __launch_bounds__(kSweepThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_den_kappa(Sc sc, uint32_t pass, const float4* __restrict__ pos,
                                                             const float4* __restrict__ vel, SolveBufs b) {
    extern __shared__ float4 smem[];
    den_kappa_on<Sc, DIV>(sc, tile_index(sc), pass, pos, vel, b, smem);
}

// Reads the working set, writes the resting set at the same slot.
template <class Sc>
__global__ __launch_bounds__(kTile) void k_integrate(Sc sc, Records work, Records rest, SolveBufs b) {
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);  // one tile a workgroup
    const SceneK& k = sc.scene(tile);
    const StepK& s = sc.step(tile);
    const uint32_t i = tile * kTile + threadIdx.x;
    const NextCount& nx = b.next;
    // The next build's look-back words and ticket, cleared as grid_count
    // would (the ones of this substep's scan are read by now).
    if (nx.cellid) {
        if (i < nx.scan_tiles) nx.status[i] = 0;
        if (i == 0) *nx.ticket = 0;
    }
    if (!sc.runs(tile, kEveryPass)) return;
    bool clamped_speed = false;
    uint32_t cell = 0xffffffffu;
    if (i < k.slot_end) {
        // Every load first: a load issued after a store would wait for it.
        const float4 was = work.vel[i], pv = work.prev_vel[i], p4 = work.pos[i];
        const float rho = b.density[i], pr = b.pressure[i];
        const float pp = work.prev_pressure[i], t = work.temperature[i];
        const uint32_t id = rest.ids ? work.ids[i] : 0u;
        f3 v = xyz(was);
        const float speed = sqrtf(dot3(v, v));
        const float ceiling = rho < 0.5f * k.rho0 ? 3.0f * s.v_clamp : s.v_clamp;
        if (speed > ceiling) {
            v = scl(ceiling / speed, v);
            clamped_speed = true;
        }
        const f3 dvv = sub(v, xyz(pv));
        const float acc = sqrtf(dot3(dvv, dvv)) / s.dt;
        const float in[3] = {p4.x + s.dt * v.x, p4.y + s.dt * v.y, p4.z + s.dt * v.z};
        float vv[3] = {v.x, v.y, v.z};
        float out[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float c = wmin(wmax(in[a], k.wall_lo[a]), k.wall_hi[a]);
            out[a] = c;
            if (c != in[a]) vv[a] = 0.0f;
        }
        rest.pos[i] = make_float4(out[0], out[1], out[2], 0.0f);
        if (nx.cellid) {
            // The next grid build's cell of this resting position
            // (grid_count's cell and non-finite test).
            if (!(__builtin_isfinite(out[0]) && __builtin_isfinite(out[1]) && __builtin_isfinite(out[2])))
                atomicOr(nx.nonfinite, 1u);
            const uint32_t cx = cell_coord(out[0], k.box_min[0], k.cell, k.dims[0]);
            const uint32_t cy = cell_coord(out[1], k.box_min[1], k.cell, k.dims[1]);
            const uint32_t cz = cell_coord(out[2], k.box_min[2], k.cell, k.dims[2]);
            cell = k.cell_offset + (cz * k.dims[1] + cy) * k.dims[0] + cx;
        }
        rest.vel[i] = make_float4(vv[0], vv[1], vv[2], wmix(was.w, acc, s.chase));
        rest.prev_vel[i] = make_float4(vv[0], vv[1], vv[2], wmix(pv.w, pr, s.chase));
        const float dp = pr - pp;
        rest.prev_pressure[i] = pr;
        rest.temperature[i] = t + t * EXPANSION / (rho * HEAT_CAPACITY) * dp;
        if (rest.ids) rest.ids[i] = id;
    }
    if (nx.cellid) {
        // The resting index is the slot (the resting set is written in this
        // substep's slot order); live lanes are a prefix of the wave.
        const bool live = i < k.slot_end;
        const uint32_t r = count_runs(cell, live, nx.counts);
        if (live) {
            nx.cellid[i] = cell;
            nx.rank[i] = r;
        }
    }
    // One atomic per wave for the clamp count, into one of the scene's
    // spread words (by tile).
    const uint64_t m = __ballot(clamped_speed);
    if (m && __lane_id() == (uint32_t)(__ffsll((unsigned long long)m) - 1))
        atomicAdd(&b.clamps[sc.index(tile) * kClampWords + (tile % kClampSlots) * kClampStride],
                  (uint32_t)__popcll(m));
}

// Tolerance mode, once per scene after the kappa sweep of a pass of one
// solve (`residual`, `limit`, `iterations` that solve's words): the mean
// error of the pass; a scene at or below tol (after at least min_passes)
// runs this pass's apply and no further pass.
// `host` (a single scene's early stop, else null): the scene's limit after
// the check, written to pinned host memory.
template <class Sc>
__global__ __launch_bounds__(kBlock) void k_converge(Sc sc, uint32_t scenes, uint32_t pass, float tol,
                                                     uint32_t min_passes, unsigned long long* __restrict__ residual,
                                                     uint32_t* __restrict__ limit,
                                                     uint32_t* __restrict__ iterations, unsigned long long* host,
                                                     uint32_t tag, uint32_t slots) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= scenes) return;
    if (sc.step_at(i).active != 0u && pass < limit[i]) {
        const SceneK& k = sc.scene_at(i);
        const uint32_t n = k.slot_end - k.slot_begin;
        const double mean = (double)take_resid(residual, i, slots) * (1.0 / 4294967296.0) / (double)(n ? n : 1u);
        iterations[i] = pass + 1;
        if (mean <= (double)tol && pass + 1 >= min_passes) limit[i] = pass + 1;
    }
    if (host) *host = ((unsigned long long)tag << 32) | limit[i];
}

// Tolerance mode, before a substep's solve: clears the residuals and sets
// each scene's limits to the caps (den_max 0: the density solve keeps its
// refine count, and its limit is that count).
template <class Sc>
__global__ __launch_bounds__(kBlock) void k_reset_passes(Sc sc, uint32_t scenes, uint32_t den_max, uint32_t div_max,
                                                         SolveBufs b, uint32_t* __restrict__ limits) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= scenes) return;
    clear_resid(b.residual, i, b.resid_slots);
    clear_resid(b.div_residual, i, b.resid_slots);
    limits[i] = den_max ? den_max : sc.step_at(i).refine;
    limits[scenes + i] = div_max;
    b.iterations[i] = den_max ? 0u : sc.step_at(i).refine;
    b.div_iterations[i] = div_max > 1u ? 0u : 1u;
}

// Tolerance mode, after a substep's solve: one histogram count per active
// scene at the passes each solve ran.
template <class Sc>
__global__ __launch_bounds__(kBlock) void k_tally_passes(Sc sc, uint32_t scenes, SolveBufs b) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= scenes || sc.step_at(i).active == 0u) return;
    const uint32_t den = b.iterations[i] < kMaxPasses ? b.iterations[i] : kMaxPasses;
    const uint32_t div = b.div_iterations[i] < kMaxPasses ? b.div_iterations[i] : kMaxPasses;
    atomicAdd(&b.pass_hist[den], 1u);
    atomicAdd(&b.pass_hist[kHistBins + div], 1u);
}

// Tolerance mode of a batch, work lists: each pass's sweeps take the tiles
// of the scenes still iterating, listed by k_converge_list, one workgroup a
// list entry (the workgroups past the list's length return at once), so that
// a pass costs the tiles it has. A list (list_words(scenes)
// words) holds the tile count, the listed scenes, the prefix of their tile
// counts (one more entry) and each one's first tile.
__host__ __device__ __forceinline__ uint32_t list_words(uint32_t scenes) { return 2u + (scenes + 1u) + scenes; }

// After the kappa sweep of `pass` (when `check`), k_converge's rule for every
// scene; then the list of the scenes that run pass `next`. One workgroup.
template <class Sc>
__global__ __launch_bounds__(1024) void k_converge_list(Sc sc, uint32_t scenes, uint32_t pass, uint32_t check,
                                                        float tol, uint32_t min_passes,
                                                        unsigned long long* __restrict__ residual,
                                                        uint32_t* __restrict__ limit, uint32_t* __restrict__ iterations,
                                                        uint32_t next, uint32_t* __restrict__ list,
                                                        unsigned long long* host, uint32_t tag, uint32_t slots) {
    __shared__ uint32_t s_tiles[16], s_scenes[16];
    uint32_t* const first = list + 2;
    uint32_t* const tile0 = list + 3 + scenes;
    const uint32_t w = threadIdx.x / 64u;
    uint32_t tiles_carry = 0, scenes_carry = 0;
    for (uint32_t base = 0; base < scenes; base += 1024u) {
        const uint32_t i = base + threadIdx.x;
        uint32_t cnt = 0, t0 = 0;
        if (i < scenes && sc.step_at(i).active != 0u) {
            const SceneK& k = sc.scene_at(i);
            uint32_t lim = limit[i];
            if (check && pass < lim) {
                const uint32_t n = k.slot_end - k.slot_begin;
                const double mean = (double)take_resid(residual, i, slots) * (1.0 / 4294967296.0) / (double)(n ? n : 1u);
                iterations[i] = pass + 1;
                if (mean <= (double)tol && pass + 1 >= min_passes) {
                    lim = pass + 1;
                    limit[i] = lim;
                }
            }

            if (next < lim && k.slot_end > k.slot_begin) {
                t0 = k.slot_begin / kTile;
                cnt = (k.slot_end + kTile - 1u) / kTile - t0;
            }
        }
        const uint32_t has = cnt ? 1u : 0u;
        const uint32_t ct = wave_inclusive_sum(cnt), hs = wave_inclusive_sum(has);
        if (threadIdx.x % 64u == 63u) {
            s_tiles[w] = ct;
            s_scenes[w] = hs;
        }
        __syncthreads();
        uint32_t tb = tiles_carry, sb = scenes_carry, tt = tiles_carry, st = scenes_carry;
        for (uint32_t j = 0; j < 16u; ++j) {
            if (j < w) {
                tb += s_tiles[j];
                sb += s_scenes[j];
            }
            tt += s_tiles[j];
            st += s_scenes[j];
        }
        if (has) {
            first[sb + hs - 1u] = tb + ct - cnt;
            tile0[sb + hs - 1u] = t0;
        }
        tiles_carry = tt;
        scenes_carry = st;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        list[0] = tiles_carry;
        list[1] = scenes_carry;
        first[scenes_carry] = tiles_carry;
        if (host) *host = ((unsigned long long)tag << 32) | scenes_carry;  // the early stop: 0 ends the solve
    }
}

// The list entry of workgroup b of a launch of g workgroups (one a tile, a
// multiple of the XCD count) over a list of `total` tiles: each XCD takes a
// contiguous run of the list, as xcd_tile does for a launch of every tile,
// and the workgroups past the list's length return at once; `flip` walks the
// list backwards.
__device__ __forceinline__ bool list_entry(uint32_t b, uint32_t g, uint32_t total, bool flip, uint32_t& e) {
    (void)g;
    const uint32_t x = b % kXcds, w = b / kXcds;
    const uint32_t hi = (uint32_t)((uint64_t)total * (x + 1) / kXcds);
    const uint32_t idx = (uint32_t)((uint64_t)total * x / kXcds) + w;
    if (idx >= hi) return false;
    e = flip ? total - 1u - idx : idx;
    return true;
}

// The tile of list entry e: the listed scene whose prefix holds e.
__device__ __forceinline__ uint32_t list_tile(const uint32_t* __restrict__ list, uint32_t scenes, uint32_t e) {
    const uint32_t* first = list + 2;
    uint32_t lo = 0, hi = list[1];
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) / 2u;
        if (first[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    return list[3 + scenes + lo] + (e - first[lo]);
}

template <class Sc, bool DIV>
__global__ __launch_bounds__(kSweepThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_den_kappa_list(
    Sc sc, uint32_t pass, const float4* __restrict__ pos, const float4* __restrict__ vel, SolveBufs b,
    const uint32_t* __restrict__ list, uint32_t scenes) {
    extern __shared__ float4 smem[];
    uint32_t e = 0;
    if (!list_entry(blockIdx.x, gridDim.x, list[0], false, e)) return;
    den_kappa_on<Sc, DIV>(sc, list_tile(list, scenes, e), pass, pos, vel, b, smem);
}

template <class Sc>
__global__ __launch_bounds__(kSweepThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_kappa_apply_list(
    Sc sc, uint32_t pass, const float4* __restrict__ pos, float4* vel, SolveBufs b,
    const uint32_t* __restrict__ list, uint32_t scenes) {
    extern __shared__ float4 smem[];
    uint32_t e = 0;
    if (!list_entry(blockIdx.x, gridDim.x, list[0], true, e)) return;
    kappa_apply_on<Sc>(sc, list_tile(list, scenes, e), pass, pos, vel, vel, b, smem);
}

constexpr size_t kKappaEntry = sizeof(float4);                         // 16
constexpr size_t kForcesEntry = 2 * sizeof(float4) + sizeof(float2) + sizeof(float);  // 44
// A tile within one cell row of its scene stages nine rows of about kTile +
// 16 slots; 13 tiles' worth leaves room for ragged rows and keeps two 256-slot
// (four 128-slot) density and den_kappa workgroups' LDS on a CU.
#ifndef FLUID_STAGE_MAX_HALF_TILES
#define FLUID_STAGE_MAX_HALF_TILES 26
#endif
constexpr uint32_t kStageMax = FLUID_STAGE_MAX_HALF_TILES * kTile / 2;

}  // namespace

uint32_t solve_tiles(uint32_t n) { return (n + kTile - 1) / kTile; }

// Sizes the staging capacity from the device's per-workgroup LDS (less 1 KB
// for static LDS) and raises the sweeps' dynamic-LDS limit to match. The
// capacity is one for every sweep, since a staged tile's list holds LDS
// indices: at most kStageMax slots; the forces sweep (44 bytes a slot) sizes the
// budget. If the device refuses the larger limit, the capacity fits 64 KB.
size_t configure_solve(SolveBufs& b, size_t lds_bytes) {
    size_t avail = lds_bytes > 1024 ? lds_bytes - 1024 : 0;
    const void* fns[] = {reinterpret_cast<const void*>(&k_density_div<OneScene>),
                         reinterpret_cast<const void*>(&k_density_div<OneScene, true>),
                         reinterpret_cast<const void*>(&k_density_div<SceneTable, true>),
                         reinterpret_cast<const void*>(&k_kappa_apply<OneScene>),
                         reinterpret_cast<const void*>(&k_forces<OneScene>),
                         reinterpret_cast<const void*>(&k_den_kappa<OneScene>),
                         reinterpret_cast<const void*>(&k_den_kappa<OneScene, true>),
                         reinterpret_cast<const void*>(&k_density_div<SceneTable>),
                         reinterpret_cast<const void*>(&k_kappa_apply<SceneTable>),
                         reinterpret_cast<const void*>(&k_forces<SceneTable>),
                         reinterpret_cast<const void*>(&k_den_kappa<SceneTable>),
                         reinterpret_cast<const void*>(&k_den_kappa<SceneTable, true>),
                         reinterpret_cast<const void*>(&k_den_kappa_list<SceneTable, false>),
                         reinterpret_cast<const void*>(&k_den_kappa_list<SceneTable, true>),
                         reinterpret_cast<const void*>(&k_kappa_apply_list<SceneTable>),
                         reinterpret_cast<const void*>(&k_den_kappa_list<OneScene, false>),
                         reinterpret_cast<const void*>(&k_den_kappa_list<OneScene, true>),
                         reinterpret_cast<const void*>(&k_kappa_apply_list<OneScene>)};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)avail) != hipSuccess)
            avail = 65536 - 1024;
    (void)hipGetLastError();
    size_t c = avail / kForcesEntry;
    c = c > kStagePad ? c - kStagePad : 0;
    if (c > kStageMax) c = kStageMax;
    b.stage_cap = (uint32_t)(c / 64 * 64);
    return avail;
}

// The substep's region records, unless the grid build's gather wrote them;
// returns whether it launched (and so recorded `start`).
template <class Sc>
static bool tile_regions(hipStream_t st, const Sc& sc, uint32_t tiles, const uint32_t* starts, const Records& work,
                         const SolveBufs& b, hipEvent_t start) {
    if (b.regions_ready) return false;
    hipExtLaunchKernelGGL(k_tile_regions<Sc>, dim3(blocks_for(tiles)), dim3(kBlock), 0, st, start, nullptr, 0, sc,
                          starts, work.pos, b.regions, tiles, b.stage_cap);
    return true;
}

template <class Sc>
static void density_only(hipStream_t st, const Sc& sc, uint32_t n, const uint32_t* starts, const Records& work,
                         const SolveBufs& b, KernelTimer* timer) {
    if (n == 0) return;
    const uint32_t tiles = solve_tiles(n);
    const Span t = span(timer, KERNEL_DENSITY_DIV);
    const bool own = tile_regions(st, sc, tiles, starts, work, b, t.a);
    hipExtLaunchKernelGGL(k_density_div<Sc>, dim3(tiles), dim3(kSweepThreads), density_lds(b), st,
                          own ? nullptr : t.a, t.z, 0, sc, starts, work.pos, work.vel, b);
}

struct NoCheck {
    void operator()(hipStream_t, uint32_t, bool) const {}
    bool stopped(uint32_t, bool) const { return false; }
};

// The early stop (PassStop) on stream `st`: issue() gives the convergence
// check of pass `pass` its tag and slot(pass) its slot, where it writes
// either the scene's limit (one scene, k_converge) or, with `counts`, the
// number of scenes listed for the next pass (k_converge_list); stopped()
// before issuing pass `pass` waits for the check `lead` passes back and
// says whether the solve ended there.
inline uint32_t stop_lead(uint32_t tiles) { return tiles >= kStopLargeTiles ? kStopLeadLarge : kStopLead; }
// Polls a slot until its check has written it. After two seconds without
// it, the stream's state decides: still running, the wait goes on; done or
// failed with the slot unwritten, the wait ends and the caller issues the
// pass, which the limits make a no-op if it is one.
inline bool wait_slot(const PassStop& p, uint32_t k, hipStream_t st, uint32_t& value) {
    auto since = std::chrono::steady_clock::now();
    for (uint32_t i = 1;; ++i) {
        const unsigned long long v = __atomic_load_n(&p.host[k], __ATOMIC_ACQUIRE);
        if ((uint32_t)(v >> 32) == p.tag[k]) {
            value = (uint32_t)v;
            return true;
        }
        if (i % 65536u == 0u && std::chrono::steady_clock::now() - since > std::chrono::seconds(2)) {
            if (hipStreamQuery(st) == hipErrorNotReady) {
                since = std::chrono::steady_clock::now();
                continue;
            }
            const unsigned long long w = __atomic_load_n(&p.host[k], __ATOMIC_ACQUIRE);
            value = (uint32_t)w;
            return (uint32_t)(w >> 32) == p.tag[k];
        }
    }
}
struct Stopper {
    PassStop* stop;
    hipStream_t st;
    bool counts;
    uint32_t lead;
    // Where the check of pass `pass` writes (null: no early stop).
    unsigned long long* slot(uint32_t pass) const { return stop ? stop->host + pass % kStopRing : nullptr; }
    // The tag of the check of pass `pass`, issued next.
    uint32_t issue(uint32_t pass) const {
        if (!stop) return 0u;
        if (++stop->next == 0u) ++stop->next;
        stop->tag[pass % kStopRing] = stop->next;
        return stop->next;
    }
    bool stopped(uint32_t pass) const {
        if (!stop || pass < lead) return false;
        uint32_t v = 0;
        if (!wait_slot(*stop, (pass - lead) % kStopRing, st, v)) return false;
        return counts ? v == 0u : v <= pass;
    }
};
// The substep's launches. With stage buffers (the adjoint's taped forward)
// each stage's velocities and kd go to arrays of their own instead of being
// updated in place: vel[0] enters div_apply, vel[1] leaves it, vel[2] leaves
// forces_den_apply, vel[3 + r] leaves den_apply of pass r; kd[0] is the
// divergence kappa, kd[1] the warm start, kd[2 + r] den_kappa of pass r.
// With tile groups, the sweeps from density_div to the last den_apply run
// group by group (tile regions and integrate cover every tile at once), in
// turn over the extra streams when there are some and no kernel timer.
// div_passes > 1 (tolerance mode) adds up to div_passes - 1 further
// divergence passes (divergence kappa, div_apply) after div_apply; `check`
// runs after each pass's kappa (density_div's for divergence pass 0).
template <class Sc, class Check = NoCheck>
static void solve(hipStream_t st, const Sc& sc, uint32_t n, const uint32_t* starts, const Records& work,
                  const Records& rest, const SolveBufs& b, uint32_t passes, KernelTimer* timer,
                  const Check& check = Check(), const StageBufs* T = nullptr, const TileGroups& G = TileGroups(),
                  uint32_t div_passes = 1) {
    if (n == 0) return;
    const uint32_t ntiles = solve_tiles(n);
    const dim3 sweep(kSweepThreads);
    const size_t lds_kappa = stage_slots(b) * kKappaEntry;
    // Velocities from the forces sweep on go to the resting record's array
    // (dead between the grid build's gather and integrate, which finishes
    // each slot there): the forces sweep reads its input's neighbours, so
    // its output cannot be in place.
    auto V = [&](uint32_t i) { return T ? T->vel[i] : (i >= 2 ? rest.vel : work.vel); };
    auto KB = [&](uint32_t i) {
        SolveBufs c = b;
        if (T) c.kd = T->kd[i];
        return c;
    };
    const uint32_t ng = G.first && G.count ? G.count : 1u;
    // A profiled substep keeps every group on one stream, so that each
    // kernel's time is its own.
    const uint32_t extra = ng > 1 && G.fork && !timer ? G.streams : 0u;
    const Span tdd0 = span(timer, KERNEL_DENSITY_DIV);
    const bool own_regions = tile_regions(st, sc, ntiles, starts, work, b, tdd0.a);
    if (extra) {
        (void)hipEventRecord(G.fork, st);
        for (uint32_t i = 0; i < extra; ++i) (void)hipStreamWaitEvent(G.stream[i], G.fork, 0);
    }
    for (uint32_t g = 0; g < ng; ++g) {
        const uint32_t t0 = G.count ? G.first[g] : 0u, t1 = G.count ? G.first[g + 1] : ntiles;
        if (t1 <= t0) continue;
        const uint32_t lane = g % (extra + 1u);
        hipStream_t s = lane ? G.stream[lane - 1u] : st;
        const dim3 tiles(t1 - t0);
        Sc fwd = sc;
        fwd.tile0 = t0;
        Sc rev = fwd;  // the reverse-order launches
        rev.flip = 1;
        const Span tdd = g == 0 ? tdd0 : span(timer, KERNEL_DENSITY_DIV);
        if (b.div_residual)
            hipExtLaunchKernelGGL(k_density_div<Sc, true>, tiles, sweep, density_lds(b), s,
                                  g == 0 && own_regions ? nullptr : tdd.a, tdd.z, 0, fwd, starts, work.pos, V(0),
                                  KB(0));
        else
            hipExtLaunchKernelGGL(k_density_div<Sc>, tiles, sweep, density_lds(b), s,
                                  g == 0 && own_regions ? nullptr : tdd.a,
                                  tdd.z, 0, fwd, starts, work.pos, V(0), KB(0));
        if (div_passes > 1) check(s, 0, true);
        const Span td = span(timer, KERNEL_DIV_APPLY);
        hipExtLaunchKernelGGL(k_kappa_apply<Sc>, tiles, sweep, lds_kappa, s, td.a, td.z, 0, rev, kEveryPass, work.pos,
                              V(0), V(1), KB(0));
        // Tolerance mode's further divergence passes, in place on the
        // velocities div_apply left (timed with div_apply).
        for (uint32_t r = 1; r < div_passes; ++r) {
            if (check.stopped(r, true)) break;
            const Span tk = span(timer, KERNEL_DIV_APPLY);
            hipExtLaunchKernelGGL(k_den_kappa<Sc, true>, tiles, sweep, stage_slots(b) * kPvEntry, s, tk.a, tk.z, 0,
                                  fwd, kDivPass | r, work.pos, V(1), KB(0));
            check(s, r, true);
            const Span tp = span(timer, KERNEL_DIV_APPLY);
            hipExtLaunchKernelGGL(k_kappa_apply<Sc>, tiles, sweep, lds_kappa, s, tp.a, tp.z, 0, rev, kDivPass | r,
                                  work.pos, V(1), V(1), KB(0));
        }
        const Span tf = span(timer, KERNEL_FORCES);
        hipExtLaunchKernelGGL(k_forces<Sc>, tiles, sweep, stage_slots(b) * kForcesEntry, s, tf.a, tf.z, 0, fwd,
                              work.pos, V(1), work.temperature, work.prev_pressure, V(2), rest.temperature, KB(1));
        // the forces sweep ran forward, so the passes start in reverse
        for (uint32_t r = 0; r < passes; ++r) {
            if (check.stopped(r, false)) break;
            const Span tk = span(timer, KERNEL_DEN_KAPPA);
            hipExtLaunchKernelGGL(k_den_kappa<Sc>, tiles, sweep, stage_slots(b) * kPvEntry, s, tk.a, tk.z, 0, rev, r,
                                  work.pos, V(2 + r), KB(2 + r));
            check(s, r, false);
            const Span tp = span(timer, KERNEL_DEN_APPLY);
            hipExtLaunchKernelGGL(k_kappa_apply<Sc>, tiles, sweep, lds_kappa, s, tp.a, tp.z, 0, fwd, r, work.pos,
                                  V(2 + r), V(3 + r), KB(2 + r));
        }
    }
    for (uint32_t i = 0; i < extra; ++i) {
        (void)hipEventRecord(G.join[i], G.stream[i]);
        (void)hipStreamWaitEvent(st, G.join[i], 0);
    }
    Records wl = work;
    wl.vel = V(passes + 2);
    wl.temperature = rest.temperature;
    const Span ti = span(timer, KERNEL_INTEGRATE);
    hipExtLaunchKernelGGL(k_integrate<Sc>, dim3(ntiles), dim3(kTile), 0, st, ti.a, ti.z, 0, sc, wl, rest, b);
}

void launch_density_only(hipStream_t st, const OneScene& sc, uint32_t slots, const uint32_t* starts,
                         const Records& work, const SolveBufs& b, KernelTimer* timer) {
    density_only(st, sc, slots, starts, work, b, timer);
}

// The reference schedule: no residuals, no pass limits.
static SolveBufs fixed_bufs(const SolveBufs& b) {
    SolveBufs f = b;
    f.residual = nullptr;
    f.div_residual = nullptr;
    return f;
}

void launch_solve(hipStream_t st, const OneScene& sc, uint32_t slots, const uint32_t* starts, const Records& work,
                  const Records& rest, const SolveBufs& b, uint32_t passes, KernelTimer* timer) {
    OneScene one = sc;
    one.s.refine = passes;
    one.pass_limit = one.div_limit = nullptr;
    solve(st, one, slots, starts, work, rest, fixed_bufs(b), passes, timer);
}

void launch_solve(hipStream_t st, const SceneTable& sc, uint32_t slots, const uint32_t* starts,
                  const Records& work, const Records& rest, const SolveBufs& b, uint32_t passes,
                  KernelTimer* timer, const TileGroups& groups) {
    SceneTable t = sc;
    t.pass_limit = t.div_limit = nullptr;
    solve(st, t, slots, starts, work, rest, fixed_bufs(b), passes, timer, NoCheck(), nullptr, groups);
}

void launch_solve_staged(hipStream_t st, const OneScene& sc, uint32_t slots, const uint32_t* starts,
                         const Records& work, const Records& rest, const SolveBufs& b, uint32_t passes,
                         const StageBufs& T) {
    OneScene one = sc;
    one.s.refine = passes;
    one.pass_limit = one.div_limit = nullptr;
    solve(st, one, slots, starts, work, rest, fixed_bufs(b), passes, nullptr, NoCheck(), &T);
}

// Both tolerance forms: reset the limits, run the solve with a convergence
// check after each kappa sweep of a solve under tolerance, tally the passes.
// One group of tiles only: a check covers every scene, so each pass must
// have run on all of them first.
template <class Sc>
static void solve_tolerance(hipStream_t st, Sc sc, uint32_t scenes, uint32_t slots, const uint32_t* starts,
                            const Records& work, const Records& rest, const SolveBufs& b, const Tolerance& tol,
                            uint32_t passes, uint32_t* limits, KernelTimer* timer) {
    const bool den = tol.tol > 0.0f, div = tol.div_tol > 0.0f && tol.div_max > 1u;
    sc.pass_limit = limits;
    sc.div_limit = limits + scenes;
    SolveBufs t = b;
    if (!den) t.residual = nullptr;
    if (!div) t.div_residual = nullptr;
    const dim3 grid(blocks_for(scenes));
    hipLaunchKernelGGL(k_reset_passes<Sc>, grid, dim3(kBlock), 0, st, sc, scenes, den ? tol.max_passes : 0u,
                       div ? tol.div_max : 1u, b, limits);
    struct Check {
        const Sc& sc;
        uint32_t scenes;
        const Tolerance& tol;
        bool den, div;
        const SolveBufs& b;
        uint32_t* limits;
        Stopper stop;
        void operator()(hipStream_t s2, uint32_t pass, bool divergence) const {
            if (divergence ? !div : !den) return;
            const uint32_t tag = stop.issue(pass);
            hipLaunchKernelGGL(k_converge<Sc>, dim3(blocks_for(scenes)), dim3(kBlock), 0, s2, sc, scenes, pass,
                               divergence ? tol.div_tol : tol.tol, divergence ? 1u : tol.min_passes,
                               divergence ? b.div_residual : b.residual, divergence ? limits + scenes : limits,
                               divergence ? b.div_iterations : b.iterations, stop.slot(pass), tag, b.resid_slots);
        }
        bool stopped(uint32_t pass, bool divergence) const {
            return (divergence ? div : den) && stop.stopped(pass);
        }
    };
    const Check check{sc, scenes, tol, den, div, b, limits, Stopper{scenes == 1u ? tol.stop : nullptr, st, false, stop_lead(solve_tiles(slots))}};
    solve(st, sc, slots, starts, work, rest, t, den ? tol.max_passes : passes, timer, check, nullptr, TileGroups(),
          div ? tol.div_max : 1u);
    hipLaunchKernelGGL(k_tally_passes<Sc>, grid, dim3(kBlock), 0, st, sc, scenes, b);
}

// A substep solve in tolerance mode with each solve under tolerance run on
// work lists (k_converge_list, k_den_kappa_list, k_kappa_apply_list): a pass
// costs the tiles of the scenes still iterating. The launches around them
// are solve()'s; the results are its bits.
template <class Sc>
static void solve_listed(hipStream_t st, Sc sc, uint32_t scenes, uint32_t n, const uint32_t* starts,
                               const Records& work, const Records& rest, const SolveBufs& b, const Tolerance& tol,
                               uint32_t passes, uint32_t* limits, KernelTimer* timer) {
    const bool den = tol.tol > 0.0f, div = tol.div_tol > 0.0f && tol.div_max > 1u;
    sc.pass_limit = limits;
    sc.div_limit = limits + scenes;
    SolveBufs t = b;
    if (!den) t.residual = nullptr;
    if (!div) t.div_residual = nullptr;
    const dim3 grid(blocks_for(scenes));
    hipLaunchKernelGGL(k_reset_passes<Sc>, grid, dim3(kBlock), 0, st, sc, scenes, den ? tol.max_passes : 0u,
                       div ? tol.div_max : 1u, b, limits);
    if (n > 0) {
        const uint32_t ntiles = solve_tiles(n);
        const dim3 tiles(ntiles), sweep(kSweepThreads);
        // one workgroup a tile, a multiple of the XCD count
        const uint32_t lg = (ntiles + kXcds - 1) / kXcds * kXcds;
        const size_t lds_kappa = stage_slots(b) * kKappaEntry, lds_pv = stage_slots(b) * kPvEntry;
        const uint32_t words = list_words(scenes);
        uint32_t* const lists[2] = {b.tile_lists, b.tile_lists + words};
        Sc rev = sc;
        rev.flip = 1;
        // The early stop (PassStop): once no scene is listed.
        const Stopper stop{tol.stop, st, true, stop_lead(ntiles)};
        auto stopped = [&](uint32_t pass) { return stop.stopped(pass); };
        // converge pass `pass` of one solve (check) and list the scenes of pass `next`
        auto converge = [&](uint32_t pass, bool check, bool divergence, uint32_t next) {
            const uint32_t tag = check ? stop.issue(pass) : 0u;
            hipLaunchKernelGGL(k_converge_list<Sc>, dim3(1), dim3(1024), 0, st, sc, scenes, pass,
                               check ? 1u : 0u, divergence ? tol.div_tol : tol.tol, divergence ? 1u : tol.min_passes,
                               divergence ? b.div_residual : b.residual, divergence ? limits + scenes : limits,
                               divergence ? b.div_iterations : b.iterations, next, lists[next % 2u],
                               check ? stop.slot(pass) : nullptr, tag, b.resid_slots);
        };
        const Span tr = span(timer, KERNEL_DENSITY_DIV);
        const bool own = tile_regions(st, sc, ntiles, starts, work, b, tr.a);
        if (div)
            hipExtLaunchKernelGGL(k_density_div<Sc, true>, tiles, sweep, density_lds(b), st, own ? nullptr : tr.a,
                                  tr.z, 0, sc, starts, work.pos, work.vel, t);
        else
            hipExtLaunchKernelGGL(k_density_div<Sc>, tiles, sweep, density_lds(b), st, own ? nullptr : tr.a,
                                  tr.z, 0, sc, starts, work.pos, work.vel, t);
        if (div) converge(0, true, true, 1);
        const Span td = span(timer, KERNEL_DIV_APPLY);
        hipExtLaunchKernelGGL(k_kappa_apply<Sc>, tiles, sweep, lds_kappa, st, td.a, td.z, 0, rev, kEveryPass,
                              work.pos, work.vel, work.vel, t);
        for (uint32_t r = 1; div && r < tol.div_max; ++r) {
            if (stopped(r)) break;
            const Span tk = span(timer, KERNEL_DIV_APPLY);
            hipExtLaunchKernelGGL(k_den_kappa_list<Sc, true>, dim3(lg), sweep, lds_pv, st, tk.a, nullptr, 0, sc,
                                  kDivPass | r, work.pos, work.vel, t, lists[r % 2u], scenes);
            converge(r, true, true, r + 1);
            hipExtLaunchKernelGGL(k_kappa_apply_list<Sc>, dim3(lg), sweep, lds_kappa, st, nullptr, tk.z, 0, sc,
                                  kDivPass | r, work.pos, work.vel, t, lists[r % 2u], scenes);
        }
        // the forces sweep writes the velocities (and temperatures) the
        // density passes and integrate take, at the resting records' arrays
        const Span tf = span(timer, KERNEL_FORCES);
        hipExtLaunchKernelGGL(k_forces<Sc>, tiles, sweep, stage_slots(b) * kForcesEntry, st, tf.a, tf.z, 0, sc,
                              work.pos, work.vel, work.temperature, work.prev_pressure, rest.vel, rest.temperature, t);
        if (den) converge(0, false, false, 0);
        const uint32_t count = den ? tol.max_passes : passes;
        for (uint32_t r = 0; r < count; ++r) {
            if (den && stopped(r)) break;
            const Span tk = span(timer, KERNEL_DEN_KAPPA);
            const Span tp = span(timer, KERNEL_DEN_APPLY);
            if (den) {
                hipExtLaunchKernelGGL(k_den_kappa_list<Sc, false>, dim3(lg), sweep, lds_pv, st, tk.a, tk.z, 0, sc, r,
                                      work.pos, rest.vel, t, lists[r % 2u], scenes);
                converge(r, true, false, r + 1);
                hipExtLaunchKernelGGL(k_kappa_apply_list<Sc>, dim3(lg), sweep, lds_kappa, st, tp.a, tp.z, 0, sc, r,
                                      work.pos, rest.vel, t, lists[r % 2u], scenes);
            } else {
                hipExtLaunchKernelGGL(k_den_kappa<Sc>, tiles, sweep, lds_pv, st, tk.a, tk.z, 0, rev, r,
                                      work.pos, rest.vel, t);
                hipExtLaunchKernelGGL(k_kappa_apply<Sc>, tiles, sweep, lds_kappa, st, tp.a, tp.z, 0, sc, r,
                                      work.pos, rest.vel, rest.vel, t);
            }
        }
        Records wl = work;
        wl.vel = rest.vel;
        wl.temperature = rest.temperature;
        const Span ti = span(timer, KERNEL_INTEGRATE);
        hipExtLaunchKernelGGL(k_integrate<Sc>, tiles, dim3(kTile), 0, st, ti.a, ti.z, 0, sc, wl, rest, t);
    }
    hipLaunchKernelGGL(k_tally_passes<Sc>, grid, dim3(kBlock), 0, st, sc, scenes, b);
}

void launch_solve_tolerance(hipStream_t st, const SceneTable& sc, uint32_t scenes, uint32_t slots,
                            const uint32_t* starts, const Records& work, const Records& rest, const SolveBufs& b,
                            const Tolerance& tol, uint32_t passes, uint32_t* limits, KernelTimer* timer) {
    if (tol.persistent && b.tile_lists)
        solve_listed(st, sc, scenes, slots, starts, work, rest, b, tol, passes, limits, timer);
    else
        solve_tolerance(st, sc, scenes, slots, starts, work, rest, b, tol, passes, limits, timer);
}

void launch_solve_tolerance(hipStream_t st, const OneScene& sc, uint32_t slots, const uint32_t* starts,
                            const Records& work, const Records& rest, const SolveBufs& b, const Tolerance& tol,
                            uint32_t passes, uint32_t* limits, KernelTimer* timer) {
    OneScene one = sc;
    one.s.refine = passes;
    if (tol.persistent && b.tile_lists)
        solve_listed(st, one, 1u, slots, starts, work, rest, b, tol, passes, limits, timer);
    else
        solve_tolerance(st, one, 1u, slots, starts, work, rest, b, tol, passes, limits, timer);
}

const SolveVariant& variant() {
    static const SolveVariant v = {kSolveTile,    &solve_tiles,  &configure_solve,        &launch_solve,
                                   &launch_solve, &launch_solve_tolerance, &launch_solve_tolerance,
                                   &launch_density_only};
    return v;
}

}  // namespace FLUID_TILE_NS
}  // namespace fluid
