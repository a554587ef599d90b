// The batched DFSPH step: many independent boxes of uneven particle counts
// stepped together, one launch per kernel per substep.
//
// Layout. Every scene's particles sit in one shared SoA set, scene s at slots
// [begin_s, begin_s + N_s), its range padded to a whole solve tile (256
// slots) so no workgroup spans two scenes. Its cells sit at cell_offset_s in
// one shared histogram, followed by one padding cell that the slots past its
// particles sort into: the one global scan then yields every scene's cell
// starts already offset to its own slots. A table maps each tile to its scene;
// the kernels read the scene's SimParams-derived constants and its Step from
// per-scene device tables.
//
// Schedule. A substep is the reference's dispatch order (render.rs:3168-3203)
// over all scenes at once; each scene uses its own Step (dt, gravity, ...) and
// its own number of density passes. In the reference schedule a scene runs
// its refine count; in tolerance mode it runs density passes until its mean
// density error is under the tolerance, and divergence passes until its mean
// divergence error is under the divergence tolerance. The whole substep can be captured once in a
// hipGraph and replayed: the per-scene Steps reach the device through a
// copy node from pinned host memory (two buffers, alternating, so the host
// fills one while the GPU reads the other).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "fluid_common.hpp"
#include "host_common.hpp"

using namespace fluid;

namespace {

constexpr int kRing = 2;
// Slots a scene group of the sweeps holds. Two groups run at once (one on
// each stream), and their neighbour lists and fields (~250 bytes a slot)
// fit the 256 MB memory-side cache together (measured at 256 scenes: 262K,
// 328K, 393K and 459K slots a group give 5.17, 5.07, 5.11 and 5.12 ms a
// substep; one stream at 786K, 5.27 ms).
constexpr size_t kGroupSlots = 327680;
constexpr uint32_t kGroupStreams = 2;

struct Ring {
    StepK* host = nullptr;   // pinned
    StepK* dev = nullptr;
    hipEvent_t done = nullptr;
    std::map<uint64_t, hipGraphExec_t> graphs;  // by (passes, tolerance)
};

}  // namespace

struct FluidBatch {
    int device = 0;
    hipStream_t stream = nullptr;
    // units: 64-slot units (slots / kMinSolveTile); tiles: the variant's tiles.
    uint32_t scenes = 0, slots = 0, cells = 0, cap = 0, units = 0, tiles = 0;
    // The solve of the tile size in use, and the launch choices the
    // autotuner sets: scene-group size (0: one group) and streams.
    const SolveVariant* V = nullptr;
    size_t group_slots = 0;
    uint32_t nstreams = 1;
    // Tile boundaries of the scene groups the sweeps run by (empty: one), and
    // the extra streams the groups take in turn, with the fork and join events
    // (created as a setting first needs them).
    std::vector<uint32_t> groups;
    uint32_t nextra = 0, created = 0;
    hipStream_t extra[kMaxGroupStreams - 1] = {};
    hipEvent_t fork = nullptr, join[kMaxGroupStreams - 1] = {};
    std::vector<SceneK> ks;
    std::vector<uint32_t> padded;
    SceneK* d_ks = nullptr;
    uint32_t* d_tile_scene = nullptr;
    uint32_t* d_pass_limit = nullptr;
    Ring ring[kRing];
    uint32_t ring_i = 0;
    Records rest{}, work{};
    GridBufs grid{};
    SolveBufs solve{};
    StatsBufs stats{};
    size_t lds_bytes = 0, lds_limit = 0;
    uint32_t stage_cap_max = 0;
    bool use_graph = false;
    Tolerance tol{0.0f, 1, 0, 0.0f, 1, true};
    // Tolerance mode's early stop on work lists (pinned slots).
    PassStop stop{};
    Physics phys{};
    EventTimer timer;
    DeviceAllocs mem;

    KernelTimer* t() { return timer.mask ? &timer : nullptr; }
    template <typename P>
    int alloc(P*& out, size_t count, bool zero = true) {
        return mem.alloc(out, count, zero);
    }
    SceneTable table(const Ring& r) const {
        SceneTable t{d_ks, r.dev, d_tile_scene};
        t.unit_shift = V->tile == 256 ? 2u : (V->tile == 128 ? 1u : 0u);
        return t;
    }
    void drop_graphs() {
        for (auto& r : ring) {
            for (auto& g : r.graphs) (void)hipGraphExecDestroy(g.second);
            r.graphs.clear();
        }
    }
    ~FluidBatch() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (auto& r : ring) {
            if (r.host) (void)hipHostFree(r.host);
            if (r.done) (void)hipEventDestroy(r.done);
        }
        for (uint32_t i = 0; i < created; ++i) (void)hipStreamSynchronize(extra[i]);
        if (stop.host) (void)hipHostFree(stop.host);
        mem.release();
        if (fork) (void)hipEventDestroy(fork);
        for (uint32_t i = 0; i < created; ++i) {
            if (join[i]) (void)hipEventDestroy(join[i]);
            (void)hipStreamDestroy(extra[i]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

bool valid(const FluidBatch* b) {
    if (!b) {
        set_error("null batch handle");
        return false;
    }
    return true;
}

bool valid_scene(const FluidBatch* b, uint32_t s) {
    if (!valid(b)) return false;
    if (s >= b->scenes) {
        set_error("scene %u out of range (%u scenes)", s, b->scenes);
        return false;
    }
    return true;
}

// Sets the solve variant, the scene groups (whole scenes of about
// group_slots slots, in the variant's tiles; 0: one group) and the streams
// the groups go round (1 to kMaxGroupStreams), creating extra streams and
// events as first needed. Captured graphs hold the old launches: the caller
// drops them.
int apply_layout(FluidBatch* B, const SolveVariant* V, size_t group_slots, uint32_t nstreams) {
    B->V = V;
    B->tiles = B->slots / V->tile;
    B->group_slots = group_slots;
    B->groups.clear();
    if (group_slots && B->slots > group_slots) {
        B->groups.push_back(0);
        size_t acc = 0;
        for (uint32_t s = 0; s < B->scenes; ++s) {
            acc += B->padded[s];
            if (acc >= group_slots && s + 1 < B->scenes) {
                B->groups.push_back(B->ks[s + 1].slot_begin / V->tile);
                acc = 0;
            }
        }
        B->groups.push_back(B->tiles);
    }
    nstreams = nstreams < 1 ? 1 : (nstreams > kMaxGroupStreams ? kMaxGroupStreams : nstreams);
    B->nstreams = nstreams;
    const uint32_t want = B->groups.size() > 2 ? nstreams - 1 : 0u;
    if (want && !B->fork) HIP_TRY(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
    while (B->created < want) {
        const uint32_t i = B->created;
        HIP_TRY(hipStreamCreateWithFlags(&B->extra[i], hipStreamNonBlocking));
        B->created = i + 1;  // counted as soon as it exists, so a failure below still frees it
        HIP_TRY(hipEventCreateWithFlags(&B->join[i], hipEventDisableTiming));
    }
    B->nextra = want;
    return FLUID_OK;
}

int init_batch(FluidBatch* B, uint32_t scenes, const uint8_t* params, uint32_t nbr_cap) {
    B->scenes = scenes;
    B->cap = nbr_cap ? nbr_cap : FLUID_SIM_NBR_CAP;
    if (B->cap % kListLanes != 0 || B->cap > 256) {
        set_error("neighbour cap %u is not a multiple of 4 in [4, 256]", B->cap);
        return FLUID_ERR_ARG;
    }
    uint64_t slots = 0, cells = 0;
    B->ks.resize(scenes);
    B->padded.resize(scenes);
    std::vector<uint32_t> tile_scene;
    for (uint32_t s = 0; s < scenes; ++s) {
        SimParamsBlock p;
        std::memcpy(&p, params + 64 * (size_t)s, sizeof p);
        if (p.dims[0] < 2 || p.dims[1] < 2 || p.dims[2] < 2) {
            set_error("scene %u: grid dims %u x %u x %u: each must be at least 2", s, p.dims[0], p.dims[1],
                      p.dims[2]);
            return FLUID_ERR_ARG;
        }
        if (!(p.cell > 0.0f) || !(p.h > 0.0f)) {
            set_error("scene %u: cell %g and h %g must be positive", s, p.cell, p.h);
            return FLUID_ERR_ARG;
        }
        SceneK k = make_scene(p);
        const uint64_t c = (uint64_t)p.dims[0] * p.dims[1] * p.dims[2];
        const uint64_t pad = ((uint64_t)p.count + kMaxSolveTile - 1) / kMaxSolveTile * kMaxSolveTile;
        if (slots + pad >= 0x80000000ull || cells + c + 1 >= 0xffffffffull) {
            set_error("batch exceeds 2^31 slots or 2^32 cells at scene %u", s);
            return FLUID_ERR_ARG;
        }
        k.cell_offset = (uint32_t)cells;
        k.cells = (uint32_t)c;
        k.slot_begin = (uint32_t)slots;
        k.slot_end = (uint32_t)(slots + p.count);
        B->ks[s] = k;
        B->padded[s] = (uint32_t)pad;
        for (uint64_t t = 0; t < pad / kMinSolveTile; ++t) tile_scene.push_back(s);
        slots += pad;
        cells += c + 1;
    }
    B->slots = (uint32_t)slots;
    B->cells = (uint32_t)cells;
    B->units = (uint32_t)tile_scene.size();
    if ((uint64_t)B->cap * std::max<uint64_t>(slots, kMaxSolveTile) >= (1ull << 32)) {
        set_error("neighbour list of cap %u x %llu slots exceeds 2^32", B->cap, (unsigned long long)slots);
        return FLUID_ERR_ARG;
    }
    // Groups of whole scenes of about kGroupSlots slots each (FLUID_BATCH_GROUP
    // overrides it; 0 runs the sweeps over the whole batch at once): a group's
    // list and fields stay in the memory-side cache from one sweep to the
    // next, where the whole batch's would not. kGroupStreams streams for the
    // groups (FLUID_BATCH_STREAMS overrides it, 1 to kMaxGroupStreams), all at
    // the default priority.
    B->group_slots = kGroupSlots;
    if (const char* env = getenv("FLUID_BATCH_GROUP")) B->group_slots = (size_t)strtoull(env, nullptr, 10);
    B->nstreams = kGroupStreams;
    if (const char* env = getenv("FLUID_BATCH_STREAMS")) B->nstreams = (uint32_t)strtoul(env, nullptr, 10);
    // Tolerance mode's passes on work lists (FLUID_TOL_PERSISTENT=0: every
    // tile dispatched each pass).
    if (const char* env = getenv("FLUID_TOL_PERSISTENT")) B->tol.persistent = strtol(env, nullptr, 10) != 0;
    HIP_TRY(hipGetDevice(&B->device));
    HIP_TRY(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    HIP_TRY(hipHostMalloc((void**)&B->stop.host, kStopRing * sizeof(unsigned long long), hipHostMallocCoherent));
    std::memset(B->stop.host, 0, kStopRing * sizeof(unsigned long long));
    int rc = apply_layout(B, solve_variant(kSolveTile), B->group_slots, B->nstreams);
    if (rc) return rc;
    const size_t n = B->slots, c = B->cells, S = scenes;
#define A(ptr, cnt)                            \
    if ((rc = B->alloc(ptr, cnt)) != FLUID_OK) \
        return rc;
    for (Records* r : {&B->rest, &B->work}) {
        A(r->pos, n);
        A(r->vel, n);
        A(r->prev_vel, n);
        A(r->prev_pressure, n);
        A(r->temperature, n);
        r->ids = nullptr;
    }
    A(B->grid.cellid, n);
    A(B->grid.rank, n);
    A(B->grid.counts, c);
    A(B->grid.starts, c + 1);
    A(B->grid.perm_tmp, n);
    A(B->grid.perm, n);
    A(B->grid.status, scan_tiles(B->cells) + 1);
    A(B->grid.ticket, 1);
    SolveBufs& b = B->solve;
    A(b.density, n);
    A(b.alpha, n);
    A(b.kd, n);
    A(b.pressure, n);
    A(b.rho_near, n);
    A(b.accel, n);
    A(b.xsph, n);
    A(b.wall_grad, n);
    b.cap = B->cap;
    b.stride = std::max<uint32_t>(B->slots, kMaxSolveTile);
    if ((rc = B->mem.alloc_list(b)) != FLUID_OK) return rc;
    A(b.nbr_n, n);
    A(b.regions, std::max<uint32_t>(B->units, 1));
    A(b.counters, 2 * S);
    B->grid.nonfinite = b.counters;
    A(b.clamps, (size_t)kClampWords * S);
    A(b.zeros, kZeroRows);
    A(b.residual, (size_t)S * kResidWords);
    b.resid_slots = resid_slots(S);
    A(b.div_residual, (size_t)S * kResidWords);
    A(b.iterations, S);
    A(b.div_iterations, S);
    A(b.pass_hist, 2 * kHistBins);
    A(b.tile_lists, 2 * (2 + (S + 1) + S));
    A(B->d_pass_limit, 2 * S);
    A(B->d_ks, S);
    A(B->d_tile_scene, std::max<uint32_t>(B->units, 1));
    int lds = 0;
    HIP_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, B->device));
    B->lds_limit = (size_t)lds;
    if (const char* env = getenv("FLUID_LDS_BYTES")) {
        const long v = strtol(env, nullptr, 10);
        if (v > 0 && (size_t)v < B->lds_limit) B->lds_limit = (size_t)v;
    }
    B->lds_bytes = B->V->configure(b, B->lds_limit);
    B->stage_cap_max = b.stage_cap;
    B->stats.blocks = 1024;
    A(B->stats.partials, S * 1024 * 16);
    A(B->stats.stats, S * 16);
#undef A
    for (auto& r : B->ring) {
        HIP_TRY(hipHostMalloc((void**)&r.host, std::max<size_t>(S, 1) * sizeof(StepK), hipHostMallocDefault));
        std::memset(r.host, 0, std::max<size_t>(S, 1) * sizeof(StepK));
        if ((rc = B->alloc(r.dev, S)) != FLUID_OK) return rc;
        HIP_TRY(hipEventCreateWithFlags(&r.done, hipEventDisableTiming));
    }
    HIP_TRY(hipMemcpy(B->d_ks, B->ks.data(), S * sizeof(SceneK), hipMemcpyHostToDevice));
    if (B->units)
        HIP_TRY(hipMemcpy(B->d_tile_scene, tile_scene.data(), tile_scene.size() * 4, hipMemcpyHostToDevice));
    std::vector<float> warm(std::max<size_t>(n, 1), 293.15f);
    HIP_TRY(hipMemcpy(B->rest.temperature, warm.data(), n * sizeof(float), hipMemcpyHostToDevice));
    return FLUID_OK;
}

// Whether a substep issues its launches directly, not from a graph: tolerance
// mode on work lists, which stops issuing passes once every scene has
// converged (PassStop) and so waits on the device between passes.
bool stops_early(const FluidBatch* B) { return B->tol.on() && B->tol.persistent; }

// The launches of one substep (the copy of the Step table included), on the
// batch stream; captured into a graph or issued directly.
void issue_substep(FluidBatch* B, const Ring& r, uint32_t passes) {
    (void)hipMemcpyAsync(r.dev, r.host, (size_t)B->scenes * sizeof(StepK), hipMemcpyHostToDevice, B->stream);
    const SceneTable sc = B->table(r);
    launch_grid_build(B->stream, sc, B->slots, B->cells, B->rest, B->work, B->grid, B->t());
    if (B->tol.on()) {
        Tolerance tol = B->tol;
        tol.stop = stops_early(B) ? &B->stop : nullptr;
        B->V->tolerance_table(B->stream, sc, B->scenes, B->slots, B->grid.starts, B->work, B->rest, B->solve,
                              tol, passes, B->d_pass_limit, B->t());
    } else {
        TileGroups g;
        if (!B->groups.empty()) {
            g.first = B->groups.data();
            g.count = (uint32_t)B->groups.size() - 1u;
            g.streams = B->nextra;
            g.fork = B->fork;
            for (uint32_t i = 0; i < B->nextra; ++i) {
                g.stream[i] = B->extra[i];
                g.join[i] = B->join[i];
            }
        }
        B->V->solve_table(B->stream, sc, B->slots, B->grid.starts, B->work, B->rest, B->solve, passes, B->t(), g);
    }
}

int run_substep(FluidBatch* B, const uint8_t* steps144, const uint32_t* refine) {
    Ring& r = B->ring[B->ring_i];
    // The GPU has finished reading this buffer's previous contents.
    HIP_TRY(hipEventSynchronize(r.done));
    uint32_t passes = 0, active = 0;
    for (uint32_t s = 0; s < B->scenes; ++s) {
        StepBlock sb;
        std::memcpy(&sb, steps144 + 144 * (size_t)s, sizeof sb);
        StepK k = make_step(sb, B->phys);
        k.refine = refine ? refine[s] : 0u;
        // dt = 0 encodes no solve (render.rs:3165-3167).
        k.active = (sb.dt > 0.0f && B->ks[s].slot_end > B->ks[s].slot_begin) ? 1u : 0u;
        if (k.active) {
            passes = std::max(passes, k.refine);
            ++active;
        }
        r.host[s] = k;
    }
    if (active == 0) return FLUID_OK;
    if (B->tol.tol > 0.0f) passes = B->tol.max_passes;
    if (B->use_graph && !B->t() && !stops_early(B)) {
        const uint64_t key = ((uint64_t)passes << 2) | (B->tol.tol > 0.0f ? 2u : 0u) | (B->tol.div_tol > 0.0f ? 1u : 0u);
        auto it = r.graphs.find(key);
        if (it == r.graphs.end()) {
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
            issue_substep(B, r, passes);
            HIP_TRY(hipStreamEndCapture(B->stream, &g));
            hipGraphExec_t exec = nullptr;
            const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIP_TRY(e);
            it = r.graphs.emplace(key, exec).first;
        }
        HIP_TRY(hipGraphLaunch(it->second, B->stream));
    } else {
        issue_substep(B, r, passes);
    }
    HIP_TRY(hipEventRecord(r.done, B->stream));
    B->ring_i = (B->ring_i + 1) % kRing;
    return check_launch("batch substep");
}

}  // namespace

extern "C" {

FluidBatch* fluid_batch_create(uint32_t scenes, const uint8_t* params64, uint32_t nbr_cap) {
    if (!params64 || scenes == 0) {
        set_error("a batch needs at least one SimParams block");
        return nullptr;
    }
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices < 1) {
        set_error("no HIP device");
        return nullptr;
    }
    FluidBatch* B = new FluidBatch();
    if (init_batch(B, scenes, params64, nbr_cap) != FLUID_OK) {
        delete B;
        return nullptr;
    }
    return B;
}

void fluid_batch_destroy(FluidBatch* b) { delete b; }

uint32_t fluid_batch_scene_count(const FluidBatch* b) { return b ? b->scenes : 0; }
uint32_t fluid_batch_slot_count(const FluidBatch* b) { return b ? b->slots : 0; }

int fluid_batch_scene_info(const FluidBatch* b, uint32_t scene, uint32_t* slot_begin, uint32_t* count,
                           uint32_t* cells) {
    if (!valid_scene(b, scene)) return FLUID_ERR_ARG;
    const SceneK& k = b->ks[scene];
    if (slot_begin) *slot_begin = k.slot_begin;
    if (count) *count = k.slot_end - k.slot_begin;
    if (cells) *cells = k.cells;
    return FLUID_OK;
}

int fluid_batch_upload_state(FluidBatch* b, uint32_t scene, const float* pos4, const float* vel4,
                             const float* prev_vel4, const float* prev_pressure, const float* temperature) {
    if (!valid_scene(b, scene)) return FLUID_ERR_ARG;
    const SceneK& k = b->ks[scene];
    const size_t o = k.slot_begin, n = k.slot_end - k.slot_begin;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (pos4) HIP_TRY(hipMemcpy(b->rest.pos + o, pos4, n * 16, hipMemcpyHostToDevice));
    if (vel4) HIP_TRY(hipMemcpy(b->rest.vel + o, vel4, n * 16, hipMemcpyHostToDevice));
    if (prev_vel4) HIP_TRY(hipMemcpy(b->rest.prev_vel + o, prev_vel4, n * 16, hipMemcpyHostToDevice));
    if (prev_pressure) HIP_TRY(hipMemcpy(b->rest.prev_pressure + o, prev_pressure, n * 4, hipMemcpyHostToDevice));
    if (temperature) HIP_TRY(hipMemcpy(b->rest.temperature + o, temperature, n * 4, hipMemcpyHostToDevice));
    return FLUID_OK;
}

int fluid_batch_download_state(FluidBatch* b, uint32_t scene, float* pos4, float* vel4, float* prev_vel4,
                               float* prev_pressure, float* temperature) {
    if (!valid_scene(b, scene)) return FLUID_ERR_ARG;
    const SceneK& k = b->ks[scene];
    const size_t o = k.slot_begin, n = k.slot_end - k.slot_begin;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (pos4) HIP_TRY(hipMemcpy(pos4, b->rest.pos + o, n * 16, hipMemcpyDeviceToHost));
    if (vel4) HIP_TRY(hipMemcpy(vel4, b->rest.vel + o, n * 16, hipMemcpyDeviceToHost));
    if (prev_vel4) HIP_TRY(hipMemcpy(prev_vel4, b->rest.prev_vel + o, n * 16, hipMemcpyDeviceToHost));
    if (prev_pressure) HIP_TRY(hipMemcpy(prev_pressure, b->rest.prev_pressure + o, n * 4, hipMemcpyDeviceToHost));
    if (temperature) HIP_TRY(hipMemcpy(temperature, b->rest.temperature + o, n * 4, hipMemcpyDeviceToHost));
    return FLUID_OK;
}

int fluid_batch_download_field(FluidBatch* b, uint32_t scene, uint32_t field, float* out) {
    if (!valid_scene(b, scene) || !out) return FLUID_ERR_ARG;
    const SceneK& k = b->ks[scene];
    const size_t o = k.slot_begin, n = k.slot_end - k.slot_begin;
    const float* src = nullptr;
    size_t per = 1;
    switch (field) {
        case FLUID_FIELD_DENSITY: src = b->solve.density; break;
        case FLUID_FIELD_ALPHA: src = b->solve.alpha; break;
        case FLUID_FIELD_KD: src = b->solve.kd; break;
        case FLUID_FIELD_PRESSURE: src = b->solve.pressure; break;
        case FLUID_FIELD_RHO_NEAR: src = b->solve.rho_near; break;
        case FLUID_FIELD_ACCEL: src = reinterpret_cast<const float*>(b->solve.accel); per = 4; break;
        case FLUID_FIELD_XSPH: src = reinterpret_cast<const float*>(b->solve.xsph); per = 4; break;
        case FLUID_FIELD_WALL_GRAD: src = reinterpret_cast<const float*>(b->solve.wall_grad); per = 4; break;
        case FLUID_FIELD_WORK_POS: src = reinterpret_cast<const float*>(b->work.pos); per = 4; break;
        case FLUID_FIELD_WORK_VEL: src = reinterpret_cast<const float*>(b->work.vel); per = 4; break;
        default:
            set_error("unknown field %u", field);
            return FLUID_ERR_ARG;
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemcpy(out, src + o * per, n * per * 4, hipMemcpyDeviceToHost));
    return FLUID_OK;
}

// The scene's grid in its own numbering: cell ids and starts relative to
// the scene's cells and slots, as a lone scene's grid would read.
int fluid_batch_download_grid(FluidBatch* b, uint32_t scene, uint32_t* cellid, uint32_t* counts, uint32_t* starts,
                              uint32_t* perm) {
    if (!valid_scene(b, scene)) return FLUID_ERR_ARG;
    const SceneK& k = b->ks[scene];
    const size_t o = k.slot_begin, n = k.slot_end - k.slot_begin, c = k.cells;
    HIP_TRY(hipStreamSynchronize(b->stream));
    std::vector<uint32_t> tmp(std::max(n, c + 1));
    if (cellid) {
        HIP_TRY(hipMemcpy(tmp.data(), b->grid.cellid + o, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) cellid[i] = tmp[i] - k.cell_offset;
    }
    if (starts || counts) {
        HIP_TRY(hipMemcpy(tmp.data(), b->grid.starts + k.cell_offset, (c + 1) * 4, hipMemcpyDeviceToHost));
        if (starts)
            for (size_t i = 0; i <= c; ++i) starts[i] = tmp[i] - k.slot_begin;
        if (counts)
            for (size_t i = 0; i < c; ++i) counts[i] = tmp[i + 1] - tmp[i];
    }
    if (perm) {
        HIP_TRY(hipMemcpy(tmp.data(), b->grid.perm + o, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) perm[i] = tmp[i] - k.slot_begin;
    }
    return FLUID_OK;
}

int fluid_batch_substep(FluidBatch* b, const uint8_t* steps144, const uint32_t* refine) {
    if (!valid(b) || !steps144) return FLUID_ERR_ARG;
    return run_substep(b, steps144, refine);
}

int fluid_batch_set_tolerance(FluidBatch* b, float tol, uint32_t min_passes, uint32_t max_passes) {
    if (!valid(b)) return FLUID_ERR_ARG;
    if (tol > 0.0f && (max_passes == 0 || min_passes > max_passes || max_passes > kMaxPasses)) {
        set_error("tolerance mode needs 0 < min_passes <= max_passes <= %u", kMaxPasses);
        return FLUID_ERR_ARG;
    }
    // A captured substep holds the tolerance and pass bounds as kernel
    // arguments, so graphs of the old setting are dropped.
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->drop_graphs();
    b->tol.tol = tol > 0.0f ? tol : 0.0f;
    b->tol.min_passes = std::max<uint32_t>(min_passes, 1);
    b->tol.max_passes = max_passes;
    return FLUID_OK;
}

int fluid_batch_set_div_tolerance(FluidBatch* b, float div_tol, uint32_t max_passes) {
    if (!valid(b)) return FLUID_ERR_ARG;
    if (div_tol > 0.0f && (max_passes == 0 || max_passes > kMaxPasses)) {
        set_error("divergence tolerance mode needs 0 < max_passes <= %u", kMaxPasses);
        return FLUID_ERR_ARG;
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->drop_graphs();
    b->tol.div_tol = div_tol > 0.0f ? div_tol : 0.0f;
    b->tol.div_max = div_tol > 0.0f ? max_passes : 1u;
    return FLUID_OK;
}

int fluid_batch_persistent_passes(FluidBatch* b, bool enable) {
    if (!valid(b)) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->drop_graphs();
    b->tol.persistent = enable;
    return FLUID_OK;
}

int fluid_batch_iterations(FluidBatch* b, uint32_t* out) {
    if (!valid(b) || !out) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemcpy(out, b->solve.iterations, (size_t)b->scenes * 4, hipMemcpyDeviceToHost));
    return FLUID_OK;
}

int fluid_batch_div_iterations(FluidBatch* b, uint32_t* out) {
    if (!valid(b) || !out) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemcpy(out, b->solve.div_iterations, (size_t)b->scenes * 4, hipMemcpyDeviceToHost));
    return FLUID_OK;
}

int fluid_batch_pass_histogram(FluidBatch* b, uint32_t* out, bool reset) {
    if (!valid(b)) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (out) HIP_TRY(hipMemcpy(out, b->solve.pass_hist, 2 * kHistBins * 4, hipMemcpyDeviceToHost));
    if (reset) HIP_TRY(hipMemset(b->solve.pass_hist, 0, 2 * kHistBins * 4));
    return FLUID_OK;
}

int fluid_batch_use_graph(FluidBatch* b, bool enable) {
    if (!valid(b)) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->use_graph = enable;
    if (!enable) b->drop_graphs();
    return FLUID_OK;
}

int fluid_batch_reduce_stats(FluidBatch* b) {
    if (!valid(b)) return FLUID_ERR_ARG;
    launch_reduce_stats(b->stream, b->table(b->ring[0]), b->scenes, b->solve.density, b->rest, b->solve.counters,
                        b->solve.clamps, b->stats, b->t());
    return check_launch("batch reduce_stats");
}

int fluid_batch_read_stats(FluidBatch* b, float* out) {
    if (!valid(b) || !out) return FLUID_ERR_ARG;
    std::vector<float> all((size_t)b->scenes * 16);
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemcpy(all.data(), b->stats.stats, all.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t s = 0; s < b->scenes; ++s) std::memcpy(out + 14 * (size_t)s, all.data() + 16 * (size_t)s, 56);
    return FLUID_OK;
}

int fluid_batch_counters(FluidBatch* b, uint64_t* out) {
    if (!valid(b) || !out) return FLUID_ERR_ARG;
    std::vector<uint32_t> c((size_t)b->scenes * 2), spread((size_t)b->scenes * kClampWords);
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemcpy(c.data(), b->solve.counters, c.size() * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(spread.data(), b->solve.clamps, spread.size() * 4, hipMemcpyDeviceToHost));
    for (size_t s = 0; s < b->scenes; ++s) {
        uint64_t clamped = 0;
        for (uint32_t w = 0; w < kClampSlots; ++w) clamped += spread[s * kClampWords + w * kClampStride];
        out[2 * s] = clamped;
        out[2 * s + 1] = c[2 * s + 1];
    }
    return FLUID_OK;
}

int fluid_batch_synchronize(FluidBatch* b) {
    if (!valid(b)) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(b->stream));
    return FLUID_OK;
}

int fluid_batch_profile(FluidBatch* b, uint32_t mask) {
    if (!valid(b)) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->timer.reset();
    b->timer.mask = mask;
    return FLUID_OK;
}

int fluid_batch_profile_read(FluidBatch* b, double* ms, uint64_t* launches) {
    if (!valid(b)) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->timer.collect();
    for (int k = 0; k < KERNEL_COUNT_; ++k) {
        if (ms) ms[k] = b->timer.ms[k];
        if (launches) launches[k] = b->timer.launches[k];
    }
    return FLUID_OK;
}

// Picks the launch shape by timing `trials` substeps of each candidate, from
// a saved copy of the resting set and the cumulative counters, then
// restores both: the results do not depend on the choices (a tile too big to
// stage gathers the same values from global memory; tiles, groups and
// streams only reorder independent work), only the time does. In turn, each
// step keeping the best of the one before: the tile size (64, 128, 256
// slots, each at its largest staging capacity), the staging capacity (0 and
// 4/13 .. 13/13 of the largest), then the scene groups and streams (one
// group; 196K, 328K or 459K-slot groups on 1 or 2 streams). Writes the
// chosen capacity. On failure the settings are left as they were.
int fluid_batch_autotune(FluidBatch* b, const uint8_t* steps144, const uint32_t* refine, uint32_t trials,
                         uint32_t* chosen) {
    if (!valid(b) || !steps144) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(b->stream));
    const size_t n = b->slots, S = b->scenes;
    Records save{};
    uint32_t *save_counters = nullptr, *save_clamps = nullptr;
    DeviceAllocs tmp;
    int rc = FLUID_OK;
    if ((rc = tmp.alloc(save.pos, n, false)) || (rc = tmp.alloc(save.vel, n, false)) ||
        (rc = tmp.alloc(save.prev_vel, n, false)) || (rc = tmp.alloc(save.prev_pressure, n, false)) ||
        (rc = tmp.alloc(save.temperature, n, false)) || (rc = tmp.alloc(save_counters, 2 * S, false)) ||
        (rc = tmp.alloc(save_clamps, kClampWords * S, false))) {
        tmp.release();
        return rc;
    }
    struct Shape {
        const SolveVariant* V;
        uint32_t cap;
        size_t group_slots;
        uint32_t streams;
    };
    const Shape keep{b->V, b->solve.stage_cap, b->group_slots, b->nstreams};
    const bool graph = b->use_graph;
    Shape best = keep;
    double best_t = 1e300;
    // Applies a shape: the variant's LDS attributes and largest capacity,
    // then the capacity (0: that largest), the groups and the streams.
    auto apply = [&](const Shape& sh) -> int {
        b->lds_bytes = sh.V->configure(b->solve, b->lds_limit);
        b->stage_cap_max = b->solve.stage_cap;
        if (sh.cap != ~0u) b->solve.stage_cap = std::min(sh.cap, b->stage_cap_max);
        return apply_layout(b, sh.V, sh.group_slots, sh.streams);
    };
    // A failed HIP call records its message and leaves through `done`, which
    // restores the settings and frees the copies.
#define TUNE_TRY(expr)                                                                                     \
    do {                                                                                                   \
        const hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) {                                                                            \
            set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__);          \
            rc = FLUID_ERR_HIP;                                                                            \
            goto done;                                                                                     \
        }                                                                                                  \
    } while (0)
    auto copy = [&](const Records& from, const Records& to, const uint32_t* c_from, uint32_t* c_to,
                    const uint32_t* k_from, uint32_t* k_to) -> hipError_t {
        hipError_t e = hipSuccess;
        const hipMemcpyKind d2d = hipMemcpyDeviceToDevice;
        if (e == hipSuccess) e = hipMemcpyAsync(to.pos, from.pos, n * 16, d2d, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(to.vel, from.vel, n * 16, d2d, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(to.prev_vel, from.prev_vel, n * 16, d2d, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(to.prev_pressure, from.prev_pressure, n * 4, d2d, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(to.temperature, from.temperature, n * 4, d2d, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(c_to, c_from, 2 * S * 4, d2d, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(k_to, k_from, kClampWords * S * 4, d2d, b->stream);
        return e;
    };
    b->use_graph = false;  // graphs bake the launch shape in
    b->drop_graphs();
    TUNE_TRY(copy(b->rest, save, b->solve.counters, save_counters, b->solve.clamps, save_clamps));
    for (int stage = 0; stage < 3; ++stage) {
        std::vector<Shape> cand;
        if (stage == 0) {
            for (uint32_t t : {64u, 128u, 256u}) cand.push_back({solve_variant(t), ~0u, best.group_slots, best.streams});
        } else if (stage == 1) {
            if ((rc = apply({best.V, ~0u, best.group_slots, best.streams}))) goto done;
            const uint32_t top = b->stage_cap_max;
            for (uint32_t f : {0u, 4u, 8u, 10u, 12u, 13u}) {
                const uint32_t c = top * f / 13u / 64u * 64u;
                if (cand.empty() || cand.back().cap != c) cand.push_back({best.V, c, best.group_slots, best.streams});
            }
        } else {
            cand.push_back({best.V, best.cap, 0, 1});
            for (size_t g : {196608ul, 327680ul, 458752ul})
                for (uint32_t st : {1u, 2u}) cand.push_back({best.V, best.cap, g, st});
        }
        for (const Shape& sh : cand) {
            if ((rc = apply(sh))) goto done;
            Shape got = sh;
            got.cap = b->solve.stage_cap;
            if ((rc = run_substep(b, steps144, refine))) goto done;  // warm-up
            TUNE_TRY(hipStreamSynchronize(b->stream));
            {
                const auto t0 = std::chrono::steady_clock::now();
                for (uint32_t i = 0; i < std::max<uint32_t>(trials, 1); ++i)
                    if ((rc = run_substep(b, steps144, refine))) goto done;
                TUNE_TRY(hipStreamSynchronize(b->stream));
                const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                if (dt < best_t) {
                    best_t = dt;
                    best = got;
                }
            }
            TUNE_TRY(copy(save, b->rest, save_counters, b->solve.counters, save_clamps, b->solve.clamps));
        }
        best_t = 1e300;  // each stage picks among its own candidates, timed alike
        if (stage == 0) best.cap = ~0u;
    }
#undef TUNE_TRY
done:
    (void)hipStreamSynchronize(b->stream);
    tmp.release();
    if (apply(rc == FLUID_OK ? best : keep) != FLUID_OK && rc == FLUID_OK) rc = FLUID_ERR_HIP;
    b->use_graph = graph;
    if (chosen) *chosen = b->solve.stage_cap;
    return rc;
}

int fluid_batch_get_tuning(const FluidBatch* b, FluidBatchTuning* out) {
    if (!valid(b) || !out) return FLUID_ERR_ARG;
    out->tile = b->V->tile;
    out->stage_cap = b->solve.stage_cap;
    out->group_slots = (uint32_t)b->group_slots;
    out->streams = b->nstreams;
    return FLUID_OK;
}

int fluid_batch_set_tuning(FluidBatch* b, const FluidBatchTuning* t) {
    if (!valid(b) || !t) return FLUID_ERR_ARG;
    const SolveVariant* v = solve_variant(t->tile);
    if (!v || t->streams < 1 || t->streams > kMaxGroupStreams) {
        set_error("tuning needs a tile of 64, 128 or 256 slots and 1 to %u streams", kMaxGroupStreams);
        return FLUID_ERR_ARG;
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->drop_graphs();
    b->lds_bytes = v->configure(b->solve, b->lds_limit);
    b->stage_cap_max = b->solve.stage_cap;
    b->solve.stage_cap = std::min(t->stage_cap / 64u * 64u, b->stage_cap_max);
    return apply_layout(b, v, t->group_slots, t->streams);
}

int fluid_batch_set_physics(FluidBatch* b, const float* phys5) {
    if (!valid(b) || !phys5) return FLUID_ERR_ARG;
    Physics ph;
    ph.mu = phys5[0];
    ph.k_near = phys5[1];
    ph.tension_scale = phys5[2];
    ph.adhesion_scale = phys5[3];
    ph.xsph_rate = phys5[4];
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (SceneK& k : b->ks) {
        const SceneK old = k;
        SimParamsBlock p{};
        for (int a = 0; a < 3; ++a) {
            p.box_min[a] = old.box_min[a];
            p.dims[a] = old.dims[a];
        }
        p.cell = old.cell;
        p.count = old.count;
        p.h = old.h;
        p.mass = old.mass;
        p.rho0 = old.rho0;
        p.inv_h = old.inv_h;
        p.sigma = old.sigma;
        p.sigma_h = old.sigma_h;
        p.support_sq = old.support_sq;
        k = make_scene(p, ph);
        k.cell_offset = old.cell_offset;
        k.cells = old.cells;
        k.slot_begin = old.slot_begin;
        k.slot_end = old.slot_end;
    }
    b->phys = ph;
    b->drop_graphs();
    HIP_TRY(hipMemcpy(b->d_ks, b->ks.data(), b->ks.size() * sizeof(SceneK), hipMemcpyHostToDevice));
    return FLUID_OK;
}

int fluid_batch_describe(const FluidBatch* b, char* buf, size_t size) {
    if (!valid(b) || !buf || size == 0) return FLUID_ERR_ARG;
    snprintf(buf, size,
             "{\"scenes\": %u, \"slots\": %u, \"cells\": %u, \"tiles\": %u, \"nbr_cap\": %u, \"stage_cap\": %u, "
             "\"lds_bytes\": %zu, \"graph\": %s, \"tolerance\": %g, \"max_passes\": %u, "
             "\"div_tolerance\": %g, \"div_max_passes\": %u, \"tile\": %u, \"group_slots\": %zu, "
             "\"groups\": %zu, \"streams\": %u}",
             b->scenes, b->slots, b->cells, b->tiles, b->cap, b->solve.stage_cap, b->lds_bytes,
             b->use_graph ? "true" : "false", (double)b->tol.tol, b->tol.max_passes, (double)b->tol.div_tol,
             b->tol.div_max, b->V->tile, b->group_slots, b->groups.empty() ? (size_t)1 : b->groups.size() - 1,
             b->groups.size() > 2 ? b->nextra + 1 : 1u);
    return FLUID_OK;
}

}  // extern "C"
