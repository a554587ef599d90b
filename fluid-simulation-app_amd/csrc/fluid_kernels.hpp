// Internal interface between the host-side sim (sim.cpp) and the kernel
// translation units (grid.hip, solve.hip, stats_tracers.hip).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fluid_common.hpp"

namespace fluid {

// Cell arithmetic only (the SimParams head the tracer splat reads).
struct GridGeom {
    float box_min[3];
    float cell;
    uint32_t dims[3];
};

// Slots a solve tile holds: 64, 128 or 256. The solve is built once for
// each (solve.hip compiled three times, into namespaces t64, t128 and t256),
// and a sim or batch picks one at run time (SolveVariant, below). Smaller
// tiles stage a little more halo per particle but put more independent
// workgroups on a CU. FLUID_SOLVE_TILE is the size a translation unit's
// unqualified solve functions and kSolveTile refer to (128 unless set).
#ifndef FLUID_SOLVE_TILE
#define FLUID_SOLVE_TILE 128
#endif
#if FLUID_SOLVE_TILE == 64
#define FLUID_TILE_NS t64
#elif FLUID_SOLVE_TILE == 128
#define FLUID_TILE_NS t128
#elif FLUID_SOLVE_TILE == 256
#define FLUID_TILE_NS t256
#else
#error "solve tile of 64, 128 or 256 slots"
#endif
constexpr uint32_t kSolveTile = FLUID_SOLVE_TILE;
// The smallest and largest tile: a batch pads each scene to whole
// kMaxSolveTile slots (so no tile of any size spans two scenes) and indexes
// its tile-to-scene table by kMinSolveTile-slot units.
constexpr uint32_t kMinSolveTile = 64;
constexpr uint32_t kMaxSolveTile = 256;

// Where a kernel takes its scene and substep constants from, per solve tile
// (kSolveTile consecutive slots, never shared by two scenes):
//   OneScene    the plain step: one scene, by value in the kernel arguments;
//   SceneTable  the batched step: per-scene tables in device memory and the
//               scene of each tile.
// The den_kappa / den_apply pass argument of a launch that runs whatever
// the scene's pass count (everything but the density passes).
constexpr uint32_t kEveryPass = 0xffffffffu;
// Tolerance mode's further divergence passes: kDivPass | r is divergence
// pass r >= 1 (pass 0 is density_div's kappa and div_apply, which always
// run).
constexpr uint32_t kDivPass = 0x80000000u;

// Whether a scene runs pass `pass` this substep: density pass r while r is
// below its pass limit (tolerance mode) or its refine count; divergence pass
// r while r is below its divergence limit (only tolerance mode launches
// them).
__device__ __forceinline__ bool pass_admits(uint32_t pass, uint32_t refine, const uint32_t* den_limit,
                                            const uint32_t* div_limit, uint32_t scene) {
    if (pass == kEveryPass) return true;
    if (pass & kDivPass) return div_limit && (pass & ~kDivPass) < div_limit[scene];
    return pass < (den_limit ? den_limit[scene] : refine);
}

// `flip` reverses a launch's tile order, and a launch covers the tiles from
// `tile0` on (see tile_index in solve.hip).
struct OneScene {
    SceneK k;
    StepK s;
    uint32_t flip = 0, tile0 = 0;
    // Tolerance mode: the density and divergence pass limits (device words,
    // lowered when the scene converges); null: the refine count and one
    // divergence pass.
    const uint32_t* pass_limit = nullptr;
    const uint32_t* div_limit = nullptr;
    __device__ __forceinline__ uint32_t index(uint32_t) const { return 0; }
    __device__ __forceinline__ const SceneK& scene_of_slot(uint32_t) const { return k; }
    __device__ __forceinline__ uint32_t index_of_slot(uint32_t) const { return 0; }
    __device__ __forceinline__ const SceneK& scene(uint32_t) const { return k; }
    __device__ __forceinline__ const StepK& step(uint32_t) const { return s; }
    // Whether the tile takes part in a launch of pass `pass`.
    __device__ __forceinline__ bool runs(uint32_t, uint32_t pass) const {
        return s.active != 0u && pass_admits(pass, s.refine, pass_limit, div_limit, 0);
    }
    // The same from a tile's region record (scene fields of TileRegion);
    // here the kernel arguments already hold them.
    __device__ __forceinline__ uint32_t index_at(uint32_t) const { return 0; }
    __device__ __forceinline__ const SceneK& scene_at(uint32_t) const { return k; }
    __device__ __forceinline__ const StepK& step_at(uint32_t) const { return s; }
    __device__ __forceinline__ uint32_t slot_end_at(uint32_t) const { return k.slot_end; }
    __device__ __forceinline__ bool runs_at(uint32_t, uint32_t, uint32_t, uint32_t pass) const {
        return runs(0, pass);
    }
};
struct SceneTable {
    const SceneK* ks;             // [scenes]
    const StepK* ss;              // [scenes]
    const uint32_t* tile_scene;   // [slots / kMinSolveTile]: the scene of each 64-slot unit
    // log2(tile / kMinSolveTile) of the solve variant launched: tile t
    // starts at unit t << unit_shift.
    uint32_t unit_shift = 0;
    // [scenes] density and divergence passes each scene runs this substep,
    // lowered when it converges (tolerance mode); null: the Step's refine
    // count and one divergence pass.
    const uint32_t* pass_limit = nullptr;
    const uint32_t* div_limit = nullptr;
    uint32_t flip = 0, tile0 = 0;
    __device__ __forceinline__ uint32_t index(uint32_t tile) const { return tile_scene[tile << unit_shift]; }
    __device__ __forceinline__ uint32_t index_of_slot(uint32_t slot) const { return tile_scene[slot / kMinSolveTile]; }
    __device__ __forceinline__ const SceneK& scene_of_slot(uint32_t slot) const { return ks[index_of_slot(slot)]; }
    __device__ __forceinline__ const SceneK& scene(uint32_t tile) const { return ks[index(tile)]; }
    __device__ __forceinline__ const StepK& step(uint32_t tile) const { return ss[index(tile)]; }
    __device__ __forceinline__ bool runs(uint32_t tile, uint32_t pass) const {
        const uint32_t i = index(tile);
        return ss[i].active != 0u && pass_admits(pass, ss[i].refine, pass_limit, div_limit, i);
    }
    __device__ __forceinline__ uint32_t index_at(uint32_t scene) const { return scene; }
    __device__ __forceinline__ const SceneK& scene_at(uint32_t scene) const { return ks[scene]; }
    __device__ __forceinline__ const StepK& step_at(uint32_t scene) const { return ss[scene]; }
    __device__ __forceinline__ uint32_t slot_end_at(uint32_t slot_end) const { return slot_end; }
    __device__ __forceinline__ bool runs_at(uint32_t scene, uint32_t active, uint32_t refine, uint32_t pass) const {
        return active != 0u && pass_admits(pass, refine, pass_limit, div_limit, scene);
    }
};

// The five persistent records of one particle set (resting or working),
// plus the optional identity payload carried through the sort.
struct Records {
    float4* pos;            // xyz; w = 0 at rest
    float4* vel;            // xyz; w = acceleration EMA
    float4* prev_vel;       // xyz; w = pressure EMA
    float* prev_pressure;   // last substep's pressure
    float* temperature;     // kelvin
    uint32_t* ids;          // null unless ids are tracked
};

struct GridBufs {
    uint32_t* cellid;    // [N] cell of each resting particle
    uint32_t* rank;      // [N] provisional rank from the histogram atomics
    uint32_t* counts;    // [C] histogram (cleared by the scan)
    uint32_t* starts;    // [C + 1]
    uint32_t* perm_tmp;  // [N] provisional slot -> resting index
    uint32_t* perm;      // [N] stable slot -> resting index
    uint64_t* status;    // [tiles] look-back words
    uint32_t* ticket;    // [1] tile ticket
    // Per scene, word 2 s of the solve's counters (null: not kept): set by
    // the grid build when a resting position has a non-finite coordinate
    // (such a particle sorts into the grid's guard layer, whose cells the
    // density walk then walks whole), cleared by integrate.
    uint32_t* nonfinite;
};

// The nine stencil-row slot ranges holding every neighbour of one tile of
// consecutive slots, and their offsets in the tile's LDS staging area: a
// neighbour slot j found in stencil row r sits at LDS index j + delta[r]. Built
// once per substep after the grid build. A tile whose ranges fit the staging
// capacity is staged in every sweep, and its neighbour list then holds LDS
// indices; otherwise the list holds slots and the sweeps gather globally.
struct TileRegion {
    int32_t delta[9];
    uint32_t base[9];  // LDS offset of row r's first slot (prefix of the lengths)
    uint32_t total;    // slots in the nine ranges
    uint32_t staged;   // total <= the staging capacity
    // The tile's scene and what decides whether it runs this substep, copied
    // here once a substep: a sweep reads them with its region in one go
    // instead of a chain of dependent loads through the scene tables.
    uint32_t scene, slot_end, active, refine;
    // Whether a particle of the tile has a nonzero wall gradient (set by
    // density_div): the later sweeps read wall_grad only for such tiles.
    uint32_t walls;
};

// The next substep's grid count, taken by integrate as it writes the
// resting positions (the same cells, histogram, provisional ranks and
// non-finite flag grid_count would compute from them, and the scan's
// look-back words cleared), so that build can skip grid_count. Null cellid:
// off (grid_count runs).
struct NextCount {
    uint32_t* cellid = nullptr;
    uint32_t* rank = nullptr;
    uint32_t* counts = nullptr;
    uint64_t* status = nullptr;
    uint32_t* ticket = nullptr;
    uint32_t* nonfinite = nullptr;  // word 0 of the scene's counters
    uint32_t scan_tiles = 0;
};

// The histogram add of a wave's 64 cells, live lanes a prefix of the wave:
// consecutive lanes that share a cell (a cell-sorted set gives runs of them)
// take one atomic between them, all runs of the wave at once. Returns the
// lane's provisional rank in its cell (the old count plus its offset in its
// run). Every lane of the wave takes part.
__device__ __forceinline__ uint32_t count_runs(uint32_t c, bool live, uint32_t* __restrict__ counts) {
    const uint32_t lane = __lane_id();
    const uint32_t prev = (uint32_t)__shfl((int)c, lane > 0 ? (int)lane - 1 : 0);
    const bool leader = live && (lane == 0 || prev != c);
    const uint64_t lead = __ballot(leader);
    const uint32_t nlive = (uint32_t)__popcll(__ballot(live));
    uint32_t base = 0;
    if (leader) {
        const uint64_t above = lane == 63 ? 0ull : lead & ~((2ull << lane) - 1ull);
        uint32_t end = above ? (uint32_t)(__ffsll((unsigned long long)above) - 1) : 64u;
        end = end < nlive ? end : nlive;
        base = atomicAdd(&counts[c], end - lane);
    }
    const uint64_t upto = lane == 63 ? lead : lead & ((2ull << lane) - 1ull);
    const uint32_t my_leader = upto ? 63u - (uint32_t)__clzll((long long)upto) : 0u;
    base = (uint32_t)__shfl((int)base, (int)my_leader);
    return base + (lane - my_leader);
}

// The scene-local cell of a position.
__device__ __forceinline__ uint32_t cell_of_pos(const SceneK& k, float4 p) {
    const uint32_t cx = cell_coord(p.x, k.box_min[0], k.cell, k.dims[0]);
    const uint32_t cy = cell_coord(p.y, k.box_min[1], k.cell, k.dims[1]);
    const uint32_t cz = cell_coord(p.z, k.box_min[2], k.cell, k.dims[2]);
    return (cz * k.dims[1] + cy) * k.dims[0] + cx;
}

// Solve tile t's region record from the scene-local cells of its first and
// last slot: each stencil row (dy, dz) of every cell
// in [c_first, c_last] lies in cells [c_first + off - 1, c_last + off + 1]
// of the row's offset off; starts (indexed at the scene's cell offset) give
// their slot ranges.
template <class Sc>
__device__ __forceinline__ void write_region(const Sc& sc, uint32_t t, int64_t c_first, int64_t c_last,
                                             const uint32_t* __restrict__ starts, TileRegion* __restrict__ out,
                                             uint32_t stage_cap) {
    const SceneK& k = sc.scene(t);
    const int64_t cells = k.cells;
    const uint32_t* __restrict__ sst = starts + k.cell_offset;
    const int64_t plane = (int64_t)k.dims[0] * k.dims[1];
    TileRegion R;
    uint32_t total = 0;
#pragma unroll
    for (uint32_t r = 0; r < 9; ++r) {
        const int64_t dz = (int64_t)(r / 3) - 1, dy = (int64_t)(r % 3) - 1;
        const int64_t off = dz * plane + dy * (int64_t)k.dims[0];
        int64_t lo = c_first + off - 1, hi = c_last + off + 1;
        if (lo < 0) lo = 0;
        if (hi > cells - 1) hi = cells - 1;
        uint32_t a = 0, e = 0;
        if (lo <= hi) {
            a = sst[lo];
            e = sst[hi + 1];
        }
        R.delta[r] = (int32_t)total - (int32_t)a;
        R.base[r] = total;
        total += e - a;
    }
    R.total = total;
    R.staged = total <= stage_cap ? 1u : 0u;
    const StepK& st = sc.step(t);
    R.scene = sc.index(t);
    R.slot_end = k.slot_end;
    R.active = st.active;
    R.refine = st.refine;
    R.walls = 0;
    out[t] = R;
}

// The region records the grid build's gather writes too (the solve's tile
// size and staging capacity), when it takes k_tile_regions' work.
struct RegionJob {
    TileRegion* regions = nullptr;
    uint32_t tiles = 0, tile = 0, stage_cap = 0;
};

// Per-substep scratch of the solver, all in slot (cell) order.
struct SolveBufs {
    float* density;
    float* alpha;
    float* kd;        // kappa / rho
    float* pressure;
    float* rho_near;
    float4* accel;    // xyz; w = dT/dt
    float4* xsph;
    float4* wall_grad;
    // Per pair, at the place the list layout gives (kListFixed, below): the
    // neighbour is its LDS index (16 bits) when the particle's tile is
    // staged, else its slot (32 bits, a separate array only unstaged tiles
    // touch); dW/r in its own array.
    uint16_t* nbr16;
    uint32_t* nbr32;
    float* nbrg;
    // density_div's scratch for the entries it cannot place while walking
    // (those past the fixed rows, whose place depends on the wave's other
    // counts): list_tail_cap(cap) a slot, slot-major. Shared by every list
    // of a context (the adjoint's tapes included).
    uint32_t* spill;
    uint32_t* nbr_n;  // [N]
    uint32_t cap;     // a multiple of 4
    uint32_t stride;  // >= the tiles' slot count
    // Per scene: [1] neighbour overflow pairs (cumulative); [0] the grid
    // build's non-finite flag (GridBufs::nonfinite).
    uint32_t* counters;
    // Per scene, the speed-clamp count (cumulative) as kClampSlots words
    // kClampStride apart, one per cache line, whose sum is the count: many
    // workgroups add to it each substep, and one word would serialise them.
    uint32_t* clamps;
    // kZeroRows zero float4s: what a sweep reads in place of wall_grad for a
    // tile with no particle near a wall (the same values, from lines that
    // stay in cache instead of HBM).
    const float4* zeros;
    // Staging capacity (slots) shared by every sweep.
    uint32_t stage_cap;
    // Tolerance mode: per scene, the fixed-point (2^-32) sums of the current
    // pass's density errors max(rho* - rho0, 0) / rho0 and divergence errors
    // max(Drho/Dt, 0) dt / rho0, and the passes of each solve this substep.
    // Null in the reference schedule. kResidWords each a scene, of which
    // resid_slots (below) are used.
    unsigned long long* residual;
    unsigned long long* div_residual;
    uint32_t* iterations;
    uint32_t* div_iterations;
    // Tolerance mode's pass-count histograms, [2][kHistBins] (density, then
    // divergence): one count per scene and substep at the passes it ran.
    uint32_t* pass_hist;
    // Tolerance mode of a batch: two work lists of the tiles still
    // iterating, 2 + (scenes + 1) + scenes words each (solve.hip).
    uint32_t* tile_lists;
    TileRegion* regions;  // [tiles]
    NextCount next;
    // The grid build's gather wrote this substep's region records (the
    // solve skips k_tile_regions).
    bool regions_ready = false;
    // The error words a scene uses, a power of two (resid_slots()).
    uint32_t resid_slots = 1;
};

constexpr uint32_t kZeroRows = 4096;
// Tolerance mode: at most kMaxPasses passes of either solve; the histograms
// hold pass counts 0 .. kMaxPasses.
constexpr uint32_t kMaxPasses = 256;
constexpr uint32_t kHistBins = kMaxPasses + 1;
constexpr uint32_t kClampSlots = 16;
constexpr uint32_t kClampStride = 32;
constexpr uint32_t kClampWords = kClampSlots * kClampStride;  // per scene

// Tolerance mode's per-scene error sums (SolveBufs::residual,
// div_residual), each as up to kResidSlots 64-bit words kResidStride words
// (one cache line) apart, a workgroup adding into the word of its tile: one
// word took every workgroup's atomic in turn, 37 us a den_kappa launch at
// 1M particles. A context uses `slots` of them a scene (resid_slots): all
// for one scene, fewer for many, whose tiles spread over the scenes' words
// already and whose checks read every scene's words. Integer sums, so the
// total is the same in any order.
constexpr uint32_t kResidSlots = 32;
constexpr uint32_t kResidStride = 16;
constexpr uint32_t kResidWords = kResidSlots * kResidStride;  // per scene
inline uint32_t resid_slots(uint32_t scenes) {
    uint32_t s = kResidSlots;
    while (s > 1u && s * scenes > kResidSlots) s /= 2u;
    return s;
}
__device__ __forceinline__ unsigned long long* resid_word(unsigned long long* base, uint32_t scene, uint32_t tile,
                                                          uint32_t slots) {
    return base + (size_t)scene * kResidWords + (tile & (slots - 1u)) * kResidStride;
}
// The scene's sum, its words cleared for the next pass. Every load is
// issued before any store (a load issued after a store waits for it).
__device__ __forceinline__ unsigned long long take_resid(unsigned long long* base, uint32_t scene,
                                                         uint32_t slots) {
    unsigned long long* w = base + (size_t)scene * kResidWords;
    unsigned long long v[kResidSlots];
#pragma unroll
    for (uint32_t i = 0; i < kResidSlots; ++i) v[i] = i < slots ? w[i * kResidStride] : 0ull;
    unsigned long long t = 0;
#pragma unroll
    for (uint32_t i = 0; i < kResidSlots; ++i) {
        t += v[i];
        if (i < slots) w[i * kResidStride] = 0;
    }
    return t;
}
__device__ __forceinline__ void clear_resid(unsigned long long* base, uint32_t scene, uint32_t slots) {
    unsigned long long* w = base + (size_t)scene * kResidWords;
    for (uint32_t i = 0; i < slots; ++i) w[i * kResidStride] = 0;
}

// List lanes a particle in the sweeps (entry m belongs to lane m % 4).
constexpr uint32_t kListLanes = 4;

// Neighbour-list layout (cap x stride entries; stride a multiple of 16).
// Entries m < 4 kListFixed (quad rows t = m / 4 < kListFixed) sit at fixed
// places: ((m / 4) * stride + p) * 4 + m % 4, so a sweep loads them before
// it knows the count and a wave's 64 lanes read one contiguous run a row.
// The rest are packed: the 16 slots of a wave (16w .. 16w + 15) own a block
// of 16 * list_tail_cap(cap) entries after the fixed rows, and slot p's
// entries past the fixed ones lie consecutively from the summed tail
// lengths of the wave's slots before p. Only the entries that exist take
// memory and line fetches.
#ifndef FLUID_LIST_FIXED
#define FLUID_LIST_FIXED 6
#endif
constexpr uint32_t kListFixed = FLUID_LIST_FIXED;
constexpr uint32_t kListWave = 64 / kListLanes;  // slots a wave
__host__ __device__ __forceinline__ uint32_t list_tail_cap(uint32_t cap) {
    return cap > kListFixed * kListLanes ? cap - kListFixed * kListLanes : 0u;
}
__host__ __device__ __forceinline__ uint32_t list_tail_len(uint32_t nn) {
    return nn > kListFixed * kListLanes ? nn - kListFixed * kListLanes : 0u;
}
// The entry of a fixed-row slot: m < 4 kListFixed.
__host__ __device__ __forceinline__ uint32_t list_fixed(uint32_t stride, uint32_t p, uint32_t m) {
    return ((m / kListLanes) * stride + p) * kListLanes + m % kListLanes;
}
// Entries each list array holds past cap x stride: a sweep's tail loads run
// up to three quads past a lane's last entry, which for the last wave's
// block is past the list.
constexpr uint32_t kListPad = 64;
// The first tail entry of slot p, whose tail starts `off` entries into its
// wave's block.
__host__ __device__ __forceinline__ uint32_t list_tail(uint32_t stride, uint32_t cap, uint32_t p, uint32_t off) {
    return kListFixed * kListLanes * stride + (p / kListWave) * kListWave * list_tail_cap(cap) + off;
}

enum KernelId {
    KERNEL_GRID_COUNT = 0,
    KERNEL_GRID_SCAN,
    KERNEL_GRID_SCATTER,
    KERNEL_DENSITY_DIV,
    KERNEL_DIV_APPLY,
    KERNEL_FORCES,        // forces_eval and forces_den_apply, one sweep
    KERNEL_FORCES_UNUSED,  // no launches (the place forces_den_apply had)
    KERNEL_DEN_KAPPA,
    KERNEL_DEN_APPLY,
    KERNEL_INTEGRATE,
    KERNEL_REDUCE_STATS,
    KERNEL_TRACERS,
    KERNEL_COUNT_
};

// HIP-event timing of kernel classes on the launch stream (profiling mode).
// begin/end record events between launches; start/stop hand out the events
// that hipExtLaunchKernelGGL records with a dispatch itself (start with a
// class's first launch, stop with its last), so no event packet sits
// between the kernels of a substep. Null when the class is not timed.
struct KernelTimer {
    virtual ~KernelTimer() = default;
    virtual void begin(hipStream_t st, int kernel) = 0;
    virtual void end(hipStream_t st, int kernel) = 0;
    virtual hipEvent_t start(int kernel) = 0;
    virtual hipEvent_t stop(int kernel) = 0;
};
struct Span {
    hipEvent_t a = nullptr, z = nullptr;
};
inline Span span(KernelTimer* t, int kernel) {
    Span s;
    if (t) {
        s.a = t->start(kernel);
        s.z = t->stop(kernel);
    }
    return s;
}

uint32_t scan_tiles(uint32_t cells);
// The grid build over `slots` resting slots and `cells` histogram cells
// (a batch's cells include one padding cell per scene).
// `counted`: the last substep's integrate already took this build's count
// (NextCount), so grid_count is skipped.
void launch_grid_build(hipStream_t st, const OneScene& sc, uint32_t slots, uint32_t cells, const Records& rest,
                       const Records& work, const GridBufs& b, KernelTimer* timer, bool counted = false,
                       const RegionJob& regions = RegionJob());
void launch_grid_build(hipStream_t st, const SceneTable& sc, uint32_t slots, uint32_t cells, const Records& rest,
                       const Records& work, const GridBufs& b, KernelTimer* timer);
void launch_scan_only(hipStream_t st, uint32_t* counts, uint32_t* starts, uint32_t cells, uint64_t* status,
                      uint32_t* ticket);
// The counting sort of the grid build into an index list only (perm), the
// positions left where they are.
void launch_index_sort(hipStream_t st, const OneScene& sc, uint32_t n, uint32_t cells, const float4* pos,
                       const GridBufs& b);

// The stage-0 microbench (render.rs:2429-2729): a seeded slab, its index
// sort and `sweeps` brute 27-cell density sweeps through the index list,
// the first frame checked on the host.
struct Stage0Bench;
Stage0Bench* stage0_create(const float extent[2], float spacing, uint32_t sweeps);
void stage0_destroy(Stage0Bench* b);
// One frame's work (sort + sweeps), synchronised; the first call also
// validates. Returns FLUID_OK or a negative status.
int stage0_frame(Stage0Bench* b);
// The first frame's verdicts (-1 before it): the scan against a CPU scan,
// every 16th density against a CPU brute-force sum to 1e-3.
void stage0_validation(const Stage0Bench* b, int* scan_ok, int* density_ok);
// Particles, cells, and the device's density (seed order) and starts.
uint32_t stage0_count(const Stage0Bench* b);
uint32_t stage0_cells(const Stage0Bench* b);
int stage0_download(Stage0Bench* b, float* density, uint32_t* starts, float* pos4);
double stage0_last_ms(const Stage0Bench* b);

// A batch may run the sweeps from density_div to the last den_apply group
// by group of whole scenes: group g is tiles [groups[g], groups[g + 1]), and
// ngroups = 0 means one group. Scenes do not interact, so the results are
// the same; a group's working set fits the memory-side cache, which the
// whole batch's does not. With more streams (an event to fork them and one
// each to join them back), the groups go round them in turn, so one group's
// last workgroups overlap the next group's first.
constexpr uint32_t kMaxGroupStreams = 4;
struct TileGroups {
    const uint32_t* first = nullptr;
    uint32_t count = 0;
    uint32_t streams = 0;  // extra streams (0 = the caller's alone)
    hipStream_t stream[kMaxGroupStreams - 1] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kMaxGroupStreams - 1] = {};
};
// The plain solve writing each stage's velocities and kd to arrays of their
// own (the adjoint's taped forward): vel[0] enters div_apply (the working
// velocities the grid build wrote), vel[1] leaves it, vel[2] leaves
// forces_den_apply, vel[3 + r] leaves den_apply of pass r; kd[0] is the
// divergence kappa, kd[1] the warm start, kd[2 + r] den_kappa of pass r.
constexpr uint32_t kMaxStagePasses = 8;
struct StageBufs {
    float4* vel[kMaxStagePasses + 3];
    float* kd[kMaxStagePasses + 2];
};
// Tolerance mode (DFSPH's two solves iterated to their tolerances): the
// divergence solve runs density_div's kappa and div_apply, then, while the
// scene's mean divergence error max(Drho/Dt, 0) dt / rho0 of the last pass
// is above div_tol, further passes of (divergence kappa, div_apply), at most
// div_max in all; the density solve runs up to max_passes passes, a scene
// stopping after the first pass (at least min_passes) whose mean density
// error max(rho* - rho0, 0) / rho0 is <= tol. A pass whose error meets the
// tolerance still applies its correction, and is the last. tol = 0 keeps
// the refine count (`passes` for a single scene, each scene's own in a
// batch); div_tol = 0 keeps one divergence pass. `limits` holds 2 x scenes
// device words (density limits, then divergence limits); sc's pass_limit
// and div_limit are set from it, and b's residuals, iterations and pass
// histograms are used.
// A batch may run the passes of each solve under tolerance on work lists
// (`persistent`): each pass's sweeps take only the tiles of the scenes still
// iterating, the workgroups past their count returning at once.
//
// A solve issued directly (not captured into a graph) can also stop issuing
// passes once it has converged (`stop`): each pass's convergence check
// writes to pinned host memory, tagged with a number the host gave that
// check, the scene's pass limit (one scene) or the number of scenes listed
// for the next pass (work lists), and before issuing pass r the host polls
// until the check of pass r - lead has written its slot and issues nothing
// more when that limit is at most r (the count is 0). No event or stream
// query goes into the queue while it polls: each one held the queue about
// 5.6 us. The passes it leaves out are those the limits would turn into
// no-ops, so the results are the same bits; `lead` passes stay queued ahead
// of the host's wait, so the device never waits for the host: kStopLead, or
// kStopLeadLarge for launches of at least kStopLargeTiles tiles, whose passes
// outlast the host's wait and issue. The full-grid form stops a single
// scene only.
constexpr uint32_t kStopLead = 4, kStopLeadLarge = 2, kStopLargeTiles = 2048, kStopRing = 8;
struct PassStop {
    // pinned (coherent), kStopRing slots the checks write: tag << 32 | value
    unsigned long long* host;
    uint32_t next = 0;              // the last tag handed out (0: none)
    uint32_t tag[kStopRing] = {};   // the tag of the check each slot awaits
};
struct Tolerance {
    float tol;
    uint32_t min_passes, max_passes;
    float div_tol;
    uint32_t div_max;
    bool persistent;
    PassStop* stop = nullptr;
    bool on() const { return tol > 0.0f || div_tol > 0.0f; }
};

// One tile size's solve, for a choice at run time.
struct SolveVariant {
    uint32_t tile;
    uint32_t (*tiles)(uint32_t n);
    size_t (*configure)(SolveBufs& b, size_t lds_bytes);
    void (*solve_one)(hipStream_t, const OneScene&, uint32_t, const uint32_t*, const Records&, const Records&,
                      const SolveBufs&, uint32_t, KernelTimer*);
    void (*solve_table)(hipStream_t, const SceneTable&, uint32_t, const uint32_t*, const Records&,
                        const Records&, const SolveBufs&, uint32_t, KernelTimer*, const TileGroups&);
    void (*tolerance_one)(hipStream_t, const OneScene&, uint32_t, const uint32_t*, const Records&,
                          const Records&, const SolveBufs&, const Tolerance&, uint32_t, uint32_t*, KernelTimer*);
    void (*tolerance_table)(hipStream_t, const SceneTable&, uint32_t, uint32_t, const uint32_t*, const Records&,
                            const Records&, const SolveBufs&, const Tolerance&, uint32_t, uint32_t*,
                            KernelTimer*);
    void (*density_only)(hipStream_t, const OneScene&, uint32_t, const uint32_t*, const Records&,
                         const SolveBufs&, KernelTimer*);
};
// The solve of one tile size (namespaces t64, t128, t256; see kSolveTile).
#define FLUID_SOLVE_API(NS)                                                                                      \
    namespace NS {                                                                                               \
    /* Tiles of the solve (consecutive slots a workgroup owns). */                                               \
    uint32_t solve_tiles(uint32_t n);                                                                            \
    /* Sets the sweeps' staging capacity for the device's per-workgroup LDS (and raises their dynamic-LDS     \
       limit); returns the dynamic LDS budget used. */                                                           \
    size_t configure_solve(SolveBufs& b, size_t lds_bytes);                                                      \
    /* One substep's solve on the working set (after the grid build), ending with integrate writing the       \
       resting set: density_div, div_apply, forces (forces_eval + forces_den_apply), passes x (den_kappa,   \
       den_apply), integrate. A scene runs min(passes, its refine count) density passes. */                               \
    void launch_solve(hipStream_t st, const OneScene& sc, uint32_t slots, const uint32_t* starts,               \
                      const Records& work, const Records& rest, const SolveBufs& b, uint32_t passes,             \
                      KernelTimer* timer);                                                                       \
    void launch_solve(hipStream_t st, const SceneTable& sc, uint32_t slots, const uint32_t* starts,             \
                      const Records& work, const Records& rest, const SolveBufs& b, uint32_t passes,             \
                      KernelTimer* timer, const TileGroups& groups = TileGroups());                              \
    void launch_solve_staged(hipStream_t st, const OneScene& sc, uint32_t slots, const uint32_t* starts,        \
                             const Records& work, const Records& rest, const SolveBufs& b, uint32_t passes,      \
                             const StageBufs& T);                                                                \
    void launch_solve_tolerance(hipStream_t st, const SceneTable& sc, uint32_t scenes, uint32_t slots,          \
                                const uint32_t* starts, const Records& work, const Records& rest,                \
                                const SolveBufs& b, const Tolerance& tol, uint32_t passes, uint32_t* limits,     \
                                KernelTimer* timer);                                                             \
    void launch_solve_tolerance(hipStream_t st, const OneScene& sc, uint32_t slots, const uint32_t* starts,     \
                                const Records& work, const Records& rest, const SolveBufs& b,                    \
                                const Tolerance& tol, uint32_t passes, uint32_t* limits, KernelTimer* timer);    \
    /* The density sweep alone (the zero-substep path of the reference's tests). */                            \
    void launch_density_only(hipStream_t st, const OneScene& sc, uint32_t slots, const uint32_t* starts,        \
                             const Records& work, const SolveBufs& b, KernelTimer* timer);                       \
    const SolveVariant& variant();                                                                               \
    }
FLUID_SOLVE_API(t64)
FLUID_SOLVE_API(t128)
FLUID_SOLVE_API(t256)
#undef FLUID_SOLVE_API
// The translation unit's own tile size unqualified (the adjoint's, 128).
using namespace FLUID_TILE_NS;

// The variant of `tile` slots (64, 128 or 256), or null.
const SolveVariant* solve_variant(uint32_t tile);

struct StatsBufs {
    float* partials;   // [blocks][16]
    uint32_t blocks;
    float* stats;      // [14]
};
uint32_t stats_blocks(uint32_t n);
void launch_reduce_stats(hipStream_t st, float rho0, uint32_t n, const float* density, const Records& rest,
                         const uint32_t* counters, const uint32_t* clamps, const StatsBufs& b, KernelTimer* timer);
// Per scene of a batch, the same fold over the scene's own slots (so a
// scene's stats are those of the plain step): partials [scenes][1024][16],
// stats [scenes][16].
void launch_reduce_stats(hipStream_t st, const SceneTable& sc, uint32_t scenes, const float* density,
                         const Records& rest, const uint32_t* counters, const uint32_t* clamps, const StatsBufs& b,
                         KernelTimer* timer);

struct TracerBufs {
    int32_t* vel_grid;   // [C][4] 16.16 sums (cleared by resolve)
    int32_t* vel_flat;   // [C][4]
    uint32_t* tracers;   // [T][2] packed
    uint32_t count;
};
void launch_tracers(hipStream_t st, const SceneK& k, const StepK& s, uint32_t n, const Records& rest,
                    const TracerBufs& b, KernelTimer* timer);

}  // namespace fluid
