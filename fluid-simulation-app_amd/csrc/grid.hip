// Neighbour-grid build for gfx950: cell hashing with an atomic per-cell
// histogram, a single-pass decoupled-look-back exclusive scan built on wave64
// DPP and ballot, and a stable scatter of the resting set into cell-sorted
// SoA arrays.
//
// Replaces the reference's count (sim_grid.wgsl:44-50), scan_single
// (sim_scan.wgsl:97-120) and scatter (sim_grid.wgsl:52-72). The reference's
// scatter takes an atomic cursor, so its in-cell order is whatever the
// atomics gave; this implementation fixes the stable order (within a cell, particles
// keep the order of the resting set), one of the orders the reference admits.
//
// Launches (all on the caller's stream, no host synchronisation):
//   grid_count   cell id per particle; histogram via one atomic per run of
//                equal cells in a wave; a provisional rank; clears the scan's
//                tile state for the current grid build. In a batch, each scene's cells
//                sit at its offset in one shared histogram, so one scan
//                segments every scene at once.
//   grid_scan    starts[c] = sum of counts below c, starts[C] = N; clears
//                the counts for the next build.
//   grid_place   provisional slot starts[c] + rank <- particle index.
//   grid_gather  stable rank within the cell; permutation; copies the five
//                persistent records (plus ids when tracked) into slot order.
#include <mutex>

#include "fluid_common.hpp"
#include "fluid_kernels.hpp"
#include "host_common.hpp"

namespace fluid {

namespace {

constexpr int kScanItems = 8;  // 4, 16 and 32 measured slower at 358K cells
constexpr uint32_t kScanTile = kBlock * kScanItems;
constexpr uint64_t kFlagAggregate = 1ull << 32;
constexpr uint64_t kFlagInclusive = 2ull << 32;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t wave_total(uint32_t x) {
    return (uint32_t)__shfl((int)wave_inclusive_sum(x), 63);
}

__device__ __forceinline__ uint64_t load_status(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Slots past a batched scene's particles (its padding to a whole solve
// tile) count into the scene's padding cell, the one after its own cells, so
// the global scan leaves every scene's slots starting at a tile boundary.
template <class Sc>
__global__ __launch_bounds__(kBlock) void k_grid_count(Sc sc, const float4* __restrict__ pos, uint32_t n,
                                                       uint32_t* __restrict__ cellid, uint32_t* __restrict__ rank,
                                                       uint32_t* __restrict__ counts, uint64_t* __restrict__ status,
                                                       uint32_t ntiles, uint32_t* __restrict__ ticket,
                                                       uint32_t* __restrict__ nonfinite) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < ntiles) status[i] = 0;
    if (i == 0) *ticket = 0;
    bool live = i < n;
    uint32_t c = 0xffffffffu;
    // A wave's 64 slots lie in one 64-slot unit, so in one scene.
    const uint32_t first = blockIdx.x * kBlock + (threadIdx.x & ~63u);
    if (first < n) {
        const SceneK& g = sc.scene_of_slot(first);
        if (live) {
            if (i < g.slot_end) {
                float4 p = pos[i];
                if (nonfinite && !(__builtin_isfinite(p.x) && __builtin_isfinite(p.y) && __builtin_isfinite(p.z)))
                    atomicOr(&nonfinite[2 * sc.index_of_slot(first)], 1u);
                uint32_t cx = cell_coord(p.x, g.box_min[0], g.cell, g.dims[0]);
                uint32_t cy = cell_coord(p.y, g.box_min[1], g.cell, g.dims[1]);
                uint32_t cz = cell_coord(p.z, g.box_min[2], g.cell, g.dims[2]);
                c = g.cell_offset + (cz * g.dims[1] + cy) * g.dims[0] + cx;
            } else {
                c = g.cell_offset + g.cells;
            }
            cellid[i] = c;
        }
    }
    // Live lanes are a prefix of the wave.
    const uint32_t r = count_runs(c, live, counts);
    if (live) rank[i] = r;
}

// One tile of kScanTile counts per workgroup, tiles taken in ticket order so
// every predecessor of a tile is already resident: its look-back waits only
// on workgroups that publish without waiting. With `clear`, the counts are
// zeroed once read (the grid build's histogram, ready for the next build).
__global__ __launch_bounds__(kBlock) void k_grid_scan(uint32_t* __restrict__ counts, uint32_t* __restrict__ starts,
                                                      uint32_t cells, uint64_t* __restrict__ status,
                                                      uint32_t* __restrict__ ticket, bool clear) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave[kBlock / 64];
    __shared__ uint32_t s_prefix;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t tsum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        uint32_t c = base + k;
        v[k] = c < cells ? counts[c] : 0u;
        tsum += v[k];
    }
    if (clear) {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            if (base + k < cells) counts[base + k] = 0u;  // ready for the next build
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t incl = wave_inclusive_sum(tsum);
    if (lane == 63u) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wave_prefix = 0, agg = 0;
#pragma unroll
    for (uint32_t w = 0; w < kBlock / 64; ++w) {
        if (w < wave) wave_prefix += s_wave[w];
        agg += s_wave[w];
    }
    if (wave == 0) {
        uint32_t prefix = 0;
        if (tile == 0) {
            if (lane == 0) store_status(&status[0], kFlagInclusive | agg);
        } else {
            if (lane == 0) store_status(&status[tile], kFlagAggregate | agg);
            int64_t j = (int64_t)tile - 1;
            while (true) {
                int64_t t = j - (int64_t)lane;
                uint64_t w = kFlagInclusive;  // before tile 0: an inclusive zero
                if (t >= 0) {
                    do {
                        w = load_status(&status[t]);
                    } while ((w >> 32) == 0);
                }
                uint64_t incl_mask = __ballot((w >> 32) == 2u);
                uint32_t val = (uint32_t)w;
                if (incl_mask) {
                    uint32_t first = (uint32_t)(__ffsll((unsigned long long)incl_mask) - 1);
                    prefix += wave_total(lane <= first ? val : 0u);
                    break;
                }
                prefix += wave_total(val);
                j -= 64;
            }
            if (lane == 0) store_status(&status[tile], kFlagInclusive | (uint64_t)(prefix + agg));
        }
        if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();
    uint32_t run = s_prefix + wave_prefix + (incl - tsum);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        uint32_t c = base + k;
        if (c < cells) starts[c] = run;
        run += v[k];
        if (c + 1 == cells) starts[cells] = run;  // the sentinel: N
    }
}

__global__ __launch_bounds__(kBlock) void k_grid_place(uint32_t n, const uint32_t* __restrict__ cellid,
                                                       const uint32_t* __restrict__ rank,
                                                       const uint32_t* __restrict__ starts,
                                                       uint32_t* __restrict__ perm_tmp) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    // A histogram that agrees with the cells puts every slot below n; the
    // guard keeps one that does not inside the array.
    const uint32_t f = starts[cellid[i]] + rank[i];
    if (f < n) perm_tmp[f] = i;
}

// A slot's particle takes the final slot start + (particles of its cell with
// a smaller resting index): the stable counting-sort order, whatever order
// the atomics of grid_count handed out. The particle's records load first,
// in flight while the rank's chain of loads (its cell, the cell's slot range,
// the cell's entries eight at a time) completes; every store follows every
// load, so no load waits on a store. Without RECORDS only the permutation
// is written (an index sort).
//
// With a RegionJob (the plain step), the launch's first region_blocks
// workgroups write the solve's tile region records instead (k_tile_regions'
// work): a tile's first and last slot hold, after place, particles of the
// cells those slots end up with, so their cells come from perm_tmp and
// cellid, with no wait for the gathered positions.
template <bool RECORDS, class Sc>
__global__ __launch_bounds__(kBlock) void k_grid_gather(uint32_t n, const uint32_t* __restrict__ perm_tmp,
                                                        const uint32_t* __restrict__ cellid,
                                                        const uint32_t* __restrict__ starts, Records rest,
                                                        Records work, uint32_t* __restrict__ perm, Sc sc,
                                                        RegionJob rj, uint32_t region_blocks) {
    if (blockIdx.x < region_blocks) {
        const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
        if (t >= rj.tiles) return;
        const SceneK& k = sc.scene(t);
        const uint32_t s0 = t * rj.tile;
        const uint32_t s1 = s0 + rj.tile < k.slot_end ? s0 + rj.tile : k.slot_end;
        const int64_t c_first = (int64_t)cellid[perm_tmp[s0]] - k.cell_offset;
        const int64_t c_last = (int64_t)cellid[perm_tmp[s1 - 1]] - k.cell_offset;
        write_region(sc, t, c_first, c_last, starts, rj.regions, rj.stage_cap);
        return;
    }
    uint32_t s = (blockIdx.x - region_blocks) * kBlock + threadIdx.x;
    if (s >= n) return;
    const uint32_t i = perm_tmp[s];
    float4 p{}, v{}, pv{};
    float pp = 0.0f, tt = 0.0f;
    uint32_t id = 0u;
    if (RECORDS) {
        p = rest.pos[i];
        v = rest.vel[i];
        pv = rest.prev_vel[i];
        pp = rest.prev_pressure[i];
        tt = rest.temperature[i];
        id = rest.ids ? rest.ids[i] : 0u;
    }
    const uint32_t c = cellid[i];
    const uint32_t st = starts[c], en = starts[c + 1] < n ? starts[c + 1] : n;
    uint32_t r = 0;
    for (uint32_t t = st; t < en; t += 8) {
        uint32_t w[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) w[u] = perm_tmp[t + u < en ? t + u : en - 1u];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) r += (t + u < en && w[u] < i) ? 1u : 0u;
    }
    const uint32_t f = st + r;
    if (f >= n) return;  // (the same guard as place's)
    perm[f] = i;
    if (!RECORDS) return;
    work.pos[f] = p;
    work.vel[f] = v;
    work.prev_vel[f] = pv;
    work.prev_pressure[f] = pp;
    work.temperature[f] = tt;
    if (rest.ids) work.ids[f] = id;
}

}  // namespace

uint32_t scan_tiles(uint32_t cells) { return (cells + kScanTile - 1) / kScanTile; }

template <class Sc>
static void grid_build(hipStream_t st, const Sc& sc, uint32_t n, uint32_t cells, const Records& rest,
                       const Records& work, const GridBufs& b, KernelTimer* timer, bool counted,
                       const RegionJob& rj) {
    const uint32_t ntiles = scan_tiles(cells);
    uint32_t count_threads = n > ntiles + 1 ? n : ntiles + 1;
    if (!counted) {
        const Span tc = span(timer, KERNEL_GRID_COUNT);
        hipExtLaunchKernelGGL(k_grid_count<Sc>, dim3(blocks_for(count_threads)), dim3(kBlock), 0, st, tc.a, tc.z, 0,
                              sc, rest.pos, n, b.cellid, b.rank, b.counts, b.status, ntiles, b.ticket, b.nonfinite);
    }
    const Span ts = span(timer, KERNEL_GRID_SCAN);
    hipExtLaunchKernelGGL(k_grid_scan, dim3(ntiles), dim3(kBlock), 0, st, ts.a, ts.z, 0, b.counts, b.starts, cells,
                          b.status, b.ticket, true);
    if (n == 0) return;
    const Span tg = span(timer, KERNEL_GRID_SCATTER);
    hipExtLaunchKernelGGL(k_grid_place, dim3(blocks_for(n)), dim3(kBlock), 0, st, tg.a, nullptr, 0, n, b.cellid,
                          b.rank, b.starts, b.perm_tmp);
    const uint32_t rb = rj.regions ? blocks_for(rj.tiles) : 0u;
    hipExtLaunchKernelGGL(k_grid_gather<true, Sc>, dim3(rb + blocks_for(n)), dim3(kBlock), 0, st, nullptr, tg.z, 0, n,
                          b.perm_tmp, b.cellid, b.starts, rest, work, b.perm, sc, rj, rb);
}

void launch_grid_build(hipStream_t st, const OneScene& sc, uint32_t slots, uint32_t cells, const Records& rest,
                       const Records& work, const GridBufs& b, KernelTimer* timer, bool counted,
                       const RegionJob& regions) {
    grid_build(st, sc, slots, cells, rest, work, b, timer, counted, regions);
}
void launch_grid_build(hipStream_t st, const SceneTable& sc, uint32_t slots, uint32_t cells, const Records& rest,
                       const Records& work, const GridBufs& b, KernelTimer* timer) {
    grid_build(st, sc, slots, cells, rest, work, b, timer, false, RegionJob());
}

// The grid build's counting sort of `n` positions into an index list (the
// stage-0 microbench's sort): cell ids, the histogram, starts (cells + 1,
// starts[cells] = n) and perm[slot] = particle, in the stable order. Nothing
// is copied.
void launch_index_sort(hipStream_t st, const OneScene& sc, uint32_t n, uint32_t cells, const float4* pos,
                       const GridBufs& b) {
    const uint32_t ntiles = scan_tiles(cells);
    const uint32_t count_threads = n > ntiles + 1 ? n : ntiles + 1;
    hipLaunchKernelGGL(k_grid_count<OneScene>, dim3(blocks_for(count_threads)), dim3(kBlock), 0, st, sc, pos, n,
                       b.cellid, b.rank, b.counts, b.status, ntiles, b.ticket, nullptr);
    hipLaunchKernelGGL(k_grid_scan, dim3(ntiles), dim3(kBlock), 0, st, b.counts, b.starts, cells, b.status, b.ticket,
                       true);
    if (n == 0) return;
    hipLaunchKernelGGL(k_grid_place, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, b.cellid, b.rank, b.starts,
                       b.perm_tmp);
    const Records none{};
    hipLaunchKernelGGL((k_grid_gather<false, OneScene>), dim3(blocks_for(n)), dim3(kBlock), 0, st, n, b.perm_tmp,
                       b.cellid, b.starts, none, none, b.perm, sc, RegionJob(), 0u);
}

// A stand-alone exclusive scan of `cells` counts (the grid scan, exposed for
// its own tests): counts are cleared, starts gets cells + 1 entries.
void launch_scan_only(hipStream_t st, uint32_t* counts, uint32_t* starts, uint32_t cells, uint64_t* status,
                      uint32_t* ticket) {
    const uint32_t ntiles = scan_tiles(cells);
    (void)hipMemsetAsync(status, 0, (size_t)ntiles * sizeof(uint64_t), st);
    (void)hipMemsetAsync(ticket, 0, sizeof(uint32_t), st);
    hipLaunchKernelGGL(k_grid_scan, dim3(ntiles), dim3(kBlock), 0, st, counts, starts, cells, status, ticket, true);
}

// ---- the scan without allocation ----------------------------------------------
//
// The look-back words and the ticket of fluid_scan_u32_noalloc live in one
// file-scope device array (each GPU holds its own copy from the moment the
// library's code loads on it until the process ends): kNoAllocTiles status
// words, then one word whose low half is the ticket. Host code reaches it
// only through the address hipGetSymbolAddress gives for the current device.
constexpr uint32_t kNoAllocMaxCells = 1u << 24;
constexpr uint32_t kNoAllocTiles = kNoAllocMaxCells / kScanTile;  // 8192: 64 KB of status words
static_assert(kNoAllocMaxCells % kScanTile == 0, "whole scan tiles");
static __device__ __align__(256) uint64_t scratch[kNoAllocTiles + 1];

namespace {

// Zeroes the first `ntiles` look-back words and the ticket (a kernel, not a
// memset, so the call can sit inside a captured graph).
__global__ __launch_bounds__(kBlock) void k_clear_scan_state(uint64_t* __restrict__ status, uint32_t ntiles,
                                                             uint32_t* __restrict__ ticket) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < ntiles) status[i] = 0;
    if (i == 0) *ticket = 0;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint64_t* g_scratch[kMaxDevices] = {};

// The scratch array's device address on the current device, looked up once
// per device.
int scratch_on_device(uint64_t** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error("hipGetDevice failed: %s", hipGetErrorString(e));
        return FLUID_ERR_HIP;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d is past the %d devices the scan's scratch table holds", dev, kMaxDevices);
        return FLUID_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scan scratch) failed on device %d: %s", dev, hipGetErrorString(e));
            return FLUID_ERR_HIP;
        }
        g_scratch[dev] = static_cast<uint64_t*>(p);
    }
    *out = g_scratch[dev];
    return FLUID_OK;
}

}  // namespace

}  // namespace fluid

using namespace fluid;

extern "C" int fluid_scan_u32_noalloc(const uint32_t* counts, uint32_t cells, uint32_t* starts, void* stream) {
    if (!counts || !starts || cells == 0) {
        set_error("scan needs device counts, device starts and cells > 0");
        return FLUID_ERR_ARG;
    }
    if (cells > kNoAllocMaxCells) {
        set_error("fluid_scan_u32_noalloc takes at most %u cells (%u given): its scratch holds %u look-back tiles",
                  kNoAllocMaxCells, cells, kNoAllocTiles);
        return FLUID_ERR_ARG;
    }
    uint64_t* base = nullptr;
    const int rc = scratch_on_device(&base);
    if (rc != FLUID_OK) return rc;
    uint64_t* status = base;
    uint32_t* ticket = reinterpret_cast<uint32_t*>(base + kNoAllocTiles);
    const uint32_t ntiles = scan_tiles(cells);
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_clear_scan_state, dim3(blocks_for(ntiles)), dim3(kBlock), 0, st, status, ntiles, ticket);
    // The kernel reads the counts and writes only starts (no clear): the
    // caller's input is left as it was.
    hipLaunchKernelGGL(k_grid_scan, dim3(ntiles), dim3(kBlock), 0, st, const_cast<uint32_t*>(counts), starts, cells,
                       status, ticket, false);
    return check_launch("fluid_scan_u32_noalloc");
}
