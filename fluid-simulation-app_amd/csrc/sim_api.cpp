// Host side of the MI355X DFSPH step: device buffers, the substep schedule
// and the C ABI of include/fluid_sim.h.
//
// Takes the role of the reference's Sim (crates/fluid-core/src/render.rs:1382-1612
// for the state it allocates, 3164-3228 for the dispatch order). Everything
// runs on one HIP stream per handle; nothing synchronises unless a call
// returns host data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "fluid_common.hpp"
#include "host_common.hpp"

namespace fluid {

SceneK make_scene(const SimParamsBlock& p, const Physics& phys) {
    SceneK k;
    for (int a = 0; a < 3; ++a) {
        k.box_min[a] = p.box_min[a];
        k.dims[a] = p.dims[a];
        // wall_lo = box_min + cell, walls at -wall_lo (sim_solve.wgsl:237-239)
        k.wall_lo[a] = p.box_min[a] + p.cell;
        k.wall_hi[a] = -k.wall_lo[a];
    }
    k.cell = p.cell;
    k.count = p.count;
    k.h = p.h;
    k.mass = p.mass;
    k.rho0 = p.rho0;
    k.inv_h = p.inv_h;
    k.sigma = p.sigma;
    k.sigma_h = p.sigma_h;
    k.support_sq = p.support_sq;
    k.inv_h_wall = 1.0f / p.h;
    // forces_eval's hoisted cohesion constants, sim_solve.wgsl:534-537
    k.coh_c = 2.0f * p.h;
    const float c3 = k.coh_c * k.coh_c * k.coh_c;
    k.coh_k = 32.0f / (3.14159265f * c3 * c3 * c3);
    k.coh_near = c3 * c3 / 64.0f;
    // tension_gamma and adhesion_beta, sim_solve.wgsl:161-172
    const float c = 2.0f * p.h;
    k.gamma = (SIGMA_WATER * 7040.0f) / (21.0f * LATTICE_SIGMA * p.rho0 * p.rho0 * c * c) * phys.tension_scale;
    k.beta = (SIGMA_WATER * (1.0f + COS_CONTACT) * 1558.4f) / (LATTICE_SIGMA * p.rho0 * p.rho0 * c * c) *
             phys.adhesion_scale;
    // the constant factors of forces_eval's pair terms, sim_solve.wgsl:554-565
    k.visc2 = 2.0f * phys.mu;
    k.heat_cond = 2.0f * CONDUCTIVITY / HEAT_CAPACITY;
    k.heat_visc = 0.5f * phys.mu / HEAT_CAPACITY;
    k.k_near = phys.k_near;
    k.eps_h2 = 0.01f * p.h * p.h;
    k.cell_offset = 0;
    k.cells = p.dims[0] * p.dims[1] * p.dims[2];
    k.slot_begin = 0;
    k.slot_end = p.count;
    return k;
}

StepK make_step(const StepBlock& s, const Physics& phys) {
    StepK k;
    for (int a = 0; a < 3; ++a) {
        k.force[a] = s.force[a];
        k.omega[a] = s.omega[a];
        k.domega[a] = s.domega[a];
    }
    k.dt = s.dt;
    k.v_clamp = s.v_clamp;
    k.seed = s.seed;
    k.touch_r = s.touch_r;
    // The shader's touches array has five slots (sim_solve.wgsl:40).
    k.touch_n = s.touch_n < 5u ? s.touch_n : 5u;
    std::memcpy(k.touches, s.touches, sizeof k.touches);
    k.c_xsph = 1.0f - std::exp(-phys.xsph_rate * s.dt);
    k.chase = 1.0f - std::exp(-LENS_RATE * s.dt);
    k.decay = std::exp(-s.dt / T_CHARGE);
    k.dt_guard = std::max(s.dt, 1e-9f);
    k.refine = 0;
    k.active = 1;
    return k;
}

}  // namespace fluid

using namespace fluid;

struct FluidSim {
    int device = 0;
    hipStream_t stream = nullptr;
    SimParamsBlock params{};
    SceneK scene{};
    Physics phys{};
    GridGeom geom{};
    uint32_t n = 0, cells = 0, cap = 0, ntr = 0;
    Records rest{}, work{};
    GridBufs grid{};
    SolveBufs solve{};
    StatsBufs stats{};
    TracerBufs tracers{};
    bool track_ids = false;
    // The last substep's integrate took the next grid build's count from the
    // resting positions it wrote (NextCount); anything else that writes the
    // resting set, or a grid build with no integrate after it, clears this.
    bool counted = false;
    // Where integrate writes that count's cells and ranks: the grid's own
    // arrays keep the last build's (fluid_sim_download_grid), and a build
    // that takes a count swaps the two.
    uint32_t* next_cellid = nullptr;
    uint32_t* next_rank = nullptr;
    size_t lds_bytes = 0;   // the dynamic LDS budget of the sweeps
    size_t lds_limit = 0;   // the device's per-workgroup LDS (FLUID_LDS_BYTES caps it)
    // The solve of the tile size in use (128 slots unless set).
    const SolveVariant* V = nullptr;
    // Tolerance mode (off: the reference schedule) and its two pass limits.
    // A single scene dispatches every tile each pass and stops issuing
    // passes once converged (PassStop): at 1M particles its passes take
    // 50 + 39 us against 71 + 51 us on work lists.
    Tolerance tol{0.0f, 1, 0, 0.0f, 1, false};
    uint32_t* limits = nullptr;
    // The solve's early stop in tolerance mode (pinned slots).
    PassStop stop{};
    EventTimer timer;
    DeviceAllocs mem;

    KernelTimer* t() { return timer.mask ? &timer : nullptr; }

    template <typename P>
    int alloc(P*& out, size_t count, bool zero = true) {
        return mem.alloc(out, count, zero);
    }

    ~FluidSim() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (stop.host) (void)hipHostFree(stop.host);
        mem.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// The list's stride: the slots rounded up to the largest tile, so one list
// layout serves every tile size.
uint64_t list_stride(uint32_t n) {
    return std::max<uint64_t>(((uint64_t)n + kMaxSolveTile - 1) / kMaxSolveTile * kMaxSolveTile, kMaxSolveTile);
}

int init_sim(FluidSim* s, const uint8_t* params64, uint32_t ntr, uint32_t nbr_cap) {
    std::memcpy(&s->params, params64, sizeof s->params);
    const SimParamsBlock& p = s->params;
    if (p.dims[0] < 2 || p.dims[1] < 2 || p.dims[2] < 2) {
        set_error("grid dims %u x %u x %u: each must be at least 2", p.dims[0], p.dims[1], p.dims[2]);
        return FLUID_ERR_ARG;
    }
    const uint64_t cells64 = (uint64_t)p.dims[0] * p.dims[1] * p.dims[2];
    if (cells64 >= 0xffffffffull) {
        set_error("grid of %llu cells exceeds 32-bit cell ids", (unsigned long long)cells64);
        return FLUID_ERR_ARG;
    }
    if (!(p.cell > 0.0f) || !(p.h > 0.0f)) {
        set_error("cell %g and h %g must be positive", p.cell, p.h);
        return FLUID_ERR_ARG;
    }
    s->scene = make_scene(p);
    for (int a = 0; a < 3; ++a) {
        s->geom.box_min[a] = p.box_min[a];
        s->geom.dims[a] = p.dims[a];
    }
    s->geom.cell = p.cell;
    s->n = p.count;
    s->cells = (uint32_t)cells64;
    s->cap = nbr_cap ? nbr_cap : FLUID_SIM_NBR_CAP;
    if (s->cap % kListLanes != 0 || s->cap > 256) {
        // The list is laid out in quads, one entry for each of a particle's
        // four lanes.
        set_error("neighbour cap %u is not a multiple of 4 in [4, 256]", s->cap);
        return FLUID_ERR_ARG;
    }
    // The list's 32-bit offsets bound cap x slots; refused before anything is
    // allocated (the allocation repeats the check).
    // Sized for every tile size: slots rounded up to the largest tile, one
    // region per smallest tile.
    const uint64_t list_entries = (uint64_t)s->cap * list_stride(s->n);
    if (list_entries >= (1ull << 32)) {
        set_error("neighbour list of %llu entries (cap %u x %u particles) exceeds 2^32",
                  (unsigned long long)list_entries, s->cap, s->n);
        return FLUID_ERR_ARG;
    }
    s->ntr = ntr;
    HIP_TRY(hipGetDevice(&s->device));
    HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    const size_t n = s->n, c = s->cells;
    int rc = FLUID_OK;
#define A(ptr, cnt)                            \
    if ((rc = s->alloc(ptr, cnt)) != FLUID_OK) \
        return rc;
    for (Records* r : {&s->rest, &s->work}) {
        A(r->pos, n);
        A(r->vel, n);
        A(r->prev_vel, n);
        A(r->prev_pressure, n);
        A(r->temperature, n);
        r->ids = nullptr;
    }
    A(s->grid.cellid, n);
    A(s->grid.rank, n);
    A(s->next_cellid, n);
    A(s->next_rank, n);
    A(s->grid.counts, c);
    A(s->grid.starts, c + 1);
    A(s->grid.perm_tmp, n);
    A(s->grid.perm, n);
    A(s->grid.status, scan_tiles(s->cells) + 1);
    A(s->grid.ticket, 1);
    SolveBufs& b = s->solve;
    A(b.density, n);
    A(b.alpha, n);
    A(b.kd, n);
    A(b.pressure, n);
    A(b.rho_near, n);
    A(b.accel, n);
    A(b.xsph, n);
    A(b.wall_grad, n);
    b.cap = s->cap;
    b.stride = (uint32_t)list_stride(s->n);
    if ((rc = s->mem.alloc_list(b)) != FLUID_OK) return rc;
    A(b.nbr_n, n);
    A(b.regions, std::max<uint32_t>((s->n + kMinSolveTile - 1) / kMinSolveTile, 1));
    A(b.counters, 4);
    s->grid.nonfinite = b.counters;
    A(b.clamps, kClampWords);
    A(b.zeros, kZeroRows);
    A(b.residual, kResidWords);
    b.resid_slots = resid_slots(1);
    A(b.div_residual, kResidWords);
    A(b.iterations, 1);
    A(b.div_iterations, 1);
    A(b.pass_hist, 2 * kHistBins);
    A(b.tile_lists, 2 * (2 + 2 + 1));
    A(s->limits, 2);
    HIP_TRY(hipHostMalloc((void**)&s->stop.host, kStopRing * sizeof(unsigned long long), hipHostMallocCoherent));
    std::memset(s->stop.host, 0, kStopRing * sizeof(unsigned long long));
    s->tol.stop = &s->stop;
    // FLUID_TOL_PERSISTENT=1: tolerance mode's passes on work lists.
    if (const char* env = getenv("FLUID_TOL_PERSISTENT")) s->tol.persistent = strtol(env, nullptr, 10) != 0;
    int lds = 0;
    HIP_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device));
    s->lds_limit = (size_t)lds;
    // FLUID_LDS_BYTES caps the LDS a sweep workgroup may take (a tuning knob).
    if (const char* env = getenv("FLUID_LDS_BYTES")) {
        const long v = strtol(env, nullptr, 10);
        if (v > 0 && (size_t)v < s->lds_limit) s->lds_limit = (size_t)v;
    }
    s->V = solve_variant(kSolveTile);
    s->lds_bytes = s->V->configure(b, s->lds_limit);
    s->stats.blocks = stats_blocks(s->n);
    A(s->stats.partials, (size_t)s->stats.blocks * 16);
    A(s->stats.stats, 16);
    A(s->tracers.vel_grid, c * 4);
    A(s->tracers.vel_flat, c * 4);
    A(s->tracers.tracers, (size_t)std::max<uint32_t>(ntr, 1) * 2);
#undef A
    s->tracers.count = ntr;
    // The box starts at 293.15 K (render.rs:1540-1555); all else at zero.
    std::vector<float> warm(std::max<size_t>(n, 1), 293.15f);
    HIP_TRY(hipMemcpy(s->rest.temperature, warm.data(), n * sizeof(float), hipMemcpyHostToDevice));
    return FLUID_OK;
}

bool valid(const FluidSim* s) {
    if (!s) {
        set_error("null sim handle");
        return false;
    }
    return true;
}

int run_substeps(FluidSim* s, const uint8_t* step144, uint32_t nsub, uint32_t refine) {
    StepBlock sb;
    std::memcpy(&sb, step144, sizeof sb);
    // dt = 0 encodes no solve: kappa divides by dt^2 (render.rs:3165-3167).
    if (!(sb.dt > 0.0f) || nsub == 0) return FLUID_OK;
    StepK k = make_step(sb, s->phys);
    k.refine = refine;  // as the solve sets it (the region records carry it)
    const OneScene sc{s->scene, k};
    // integrate takes the next build's count when its launch covers the
    // scan's look-back words, which it clears.
    const uint32_t scan_t = scan_tiles(s->cells);
    const bool fuse = s->n > 0 && (uint64_t)s->V->tiles(s->n) * s->V->tile > scan_t;
    SolveBufs b = s->solve;
    // The gather writes the solve's tile region records.
    RegionJob rj;
    if (s->n > 0) {
        rj.regions = s->solve.regions;
        rj.tiles = s->V->tiles(s->n);
        rj.tile = s->V->tile;
        rj.stage_cap = s->solve.stage_cap;
        b.regions_ready = true;
    }
    for (uint32_t i = 0; i < nsub; ++i) {
        if (s->counted) {
            std::swap(s->grid.cellid, s->next_cellid);
            std::swap(s->grid.rank, s->next_rank);
        }
        if (fuse) {
            b.next.cellid = s->next_cellid;
            b.next.rank = s->next_rank;
            b.next.counts = s->grid.counts;
            b.next.status = s->grid.status;
            b.next.ticket = s->grid.ticket;
            b.next.nonfinite = s->grid.nonfinite;
            b.next.scan_tiles = scan_t;
        }
        launch_grid_build(s->stream, sc, s->n, s->cells, s->rest, s->work, s->grid, s->t(), s->counted, rj);
        if (s->tol.on())
            s->V->tolerance_one(s->stream, sc, s->n, s->grid.starts, s->work, s->rest, b, s->tol, refine,
                                s->limits, s->t());
        else
            s->V->solve_one(s->stream, sc, s->n, s->grid.starts, s->work, s->rest, b, refine, s->t());
        s->counted = fuse;
    }
    return check_launch("substep");
}

}  // namespace

extern "C" {

const char* fluid_sim_last_error(void) { return last_error(); }

int fluid_sim_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

FluidSim* fluid_sim_create(const uint8_t* params64, uint32_t tracers, uint32_t nbr_cap) {
    if (!params64) {
        set_error("null SimParams block");
        return nullptr;
    }
    if (fluid_sim_device_count() < 1) {
        set_error("no HIP device");
        return nullptr;
    }
    FluidSim* s = new FluidSim();
    if (init_sim(s, params64, tracers, nbr_cap) != FLUID_OK) {
        delete s;
        return nullptr;
    }
    return s;
}

void fluid_sim_destroy(FluidSim* sim) { delete sim; }

uint32_t fluid_sim_count(const FluidSim* sim) { return sim ? sim->n : 0; }
uint32_t fluid_sim_cells(const FluidSim* sim) { return sim ? sim->cells : 0; }
uint32_t fluid_sim_tracer_count(const FluidSim* sim) { return sim ? sim->ntr : 0; }

int fluid_sim_upload_state(FluidSim* s, const float* pos4, const float* vel4, const float* prev_vel4,
                           const float* prev_pressure, const float* temperature) {
    if (!valid(s)) return FLUID_ERR_ARG;
    const size_t n = s->n;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->counted) {
        // The count integrate took is of positions this upload replaces: the
        // next build counts again, into a cleared histogram.
        HIP_TRY(hipMemset(s->grid.counts, 0, (size_t)s->cells * 4));
        s->counted = false;
    }
    if (pos4) HIP_TRY(hipMemcpy(s->rest.pos, pos4, n * 16, hipMemcpyHostToDevice));
    if (vel4) HIP_TRY(hipMemcpy(s->rest.vel, vel4, n * 16, hipMemcpyHostToDevice));
    if (prev_vel4) HIP_TRY(hipMemcpy(s->rest.prev_vel, prev_vel4, n * 16, hipMemcpyHostToDevice));
    if (prev_pressure) HIP_TRY(hipMemcpy(s->rest.prev_pressure, prev_pressure, n * 4, hipMemcpyHostToDevice));
    if (temperature) HIP_TRY(hipMemcpy(s->rest.temperature, temperature, n * 4, hipMemcpyHostToDevice));
    return FLUID_OK;
}

int fluid_sim_download_state(FluidSim* s, float* pos4, float* vel4, float* prev_vel4, float* prev_pressure,
                             float* temperature) {
    if (!valid(s)) return FLUID_ERR_ARG;
    const size_t n = s->n;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (pos4) HIP_TRY(hipMemcpy(pos4, s->rest.pos, n * 16, hipMemcpyDeviceToHost));
    if (vel4) HIP_TRY(hipMemcpy(vel4, s->rest.vel, n * 16, hipMemcpyDeviceToHost));
    if (prev_vel4) HIP_TRY(hipMemcpy(prev_vel4, s->rest.prev_vel, n * 16, hipMemcpyDeviceToHost));
    if (prev_pressure) HIP_TRY(hipMemcpy(prev_pressure, s->rest.prev_pressure, n * 4, hipMemcpyDeviceToHost));
    if (temperature) HIP_TRY(hipMemcpy(temperature, s->rest.temperature, n * 4, hipMemcpyDeviceToHost));
    return FLUID_OK;
}

int fluid_sim_download_field(FluidSim* s, uint32_t field, float* out) {
    if (!valid(s) || !out) return FLUID_ERR_ARG;
    const void* src = nullptr;
    size_t per = 1;
    switch (field) {
        case FLUID_FIELD_DENSITY: src = s->solve.density; break;
        case FLUID_FIELD_ALPHA: src = s->solve.alpha; break;
        case FLUID_FIELD_KD: src = s->solve.kd; break;
        case FLUID_FIELD_PRESSURE: src = s->solve.pressure; break;
        case FLUID_FIELD_RHO_NEAR: src = s->solve.rho_near; break;
        case FLUID_FIELD_ACCEL: src = s->solve.accel; per = 4; break;
        case FLUID_FIELD_XSPH: src = s->solve.xsph; per = 4; break;
        case FLUID_FIELD_WALL_GRAD: src = s->solve.wall_grad; per = 4; break;
        case FLUID_FIELD_WORK_POS: src = s->work.pos; per = 4; break;
        case FLUID_FIELD_WORK_VEL: src = s->work.vel; per = 4; break;
        default:
            set_error("unknown field %u", field);
            return FLUID_ERR_ARG;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(out, src, (size_t)s->n * per * 4, hipMemcpyDeviceToHost));
    return FLUID_OK;
}

int fluid_sim_download_grid(FluidSim* s, uint32_t* cellid, uint32_t* counts, uint32_t* starts, uint32_t* perm,
                            uint32_t* nbr_n) {
    if (!valid(s)) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    const size_t n = s->n, c = s->cells;
    if (cellid) HIP_TRY(hipMemcpy(cellid, s->grid.cellid, n * 4, hipMemcpyDeviceToHost));
    if (starts) HIP_TRY(hipMemcpy(starts, s->grid.starts, (c + 1) * 4, hipMemcpyDeviceToHost));
    if (counts) {
        // The scan clears the histogram for the next build; the counts of the
        // last build are the differences of the starts.
        std::vector<uint32_t> st(c + 1);
        HIP_TRY(hipMemcpy(st.data(), s->grid.starts, (c + 1) * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < c; ++i) counts[i] = st[i + 1] - st[i];
    }
    if (perm) HIP_TRY(hipMemcpy(perm, s->grid.perm, n * 4, hipMemcpyDeviceToHost));
    if (nbr_n) HIP_TRY(hipMemcpy(nbr_n, s->solve.nbr_n, n * 4, hipMemcpyDeviceToHost));
    return FLUID_OK;
}

int fluid_sim_download_neighbours(FluidSim* s, uint32_t* nbr) {
    if (!valid(s) || !nbr) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    const size_t n = s->n, cap = s->cap, stride = s->solve.stride;
    if (n == 0) return FLUID_OK;
    std::vector<uint16_t> raw16(cap * stride);
    std::vector<uint32_t> raw32(cap * stride);
    std::vector<uint32_t> cnt(n);
    std::vector<TileRegion> reg(s->V->tiles(s->n));
    HIP_TRY(hipMemcpy(raw16.data(), s->solve.nbr16, raw16.size() * 2, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(raw32.data(), s->solve.nbr32, raw32.size() * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cnt.data(), s->solve.nbr_n, n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(reg.data(), s->solve.regions, reg.size() * sizeof(TileRegion), hipMemcpyDeviceToHost));
    uint32_t off = 0;  // the slot's tail offset in its wave's block
    for (size_t i = 0; i < n; ++i) {
        if (i % kListWave == 0) off = 0;
        const TileRegion& R = reg[i / s->V->tile];
        for (size_t m = 0; m < cap; ++m) {
            uint32_t j = 0;
            if (m < cnt[i]) {
                const size_t e = m < kListFixed * kListLanes
                                     ? list_fixed((uint32_t)stride, (uint32_t)i, (uint32_t)m)
                                     : list_tail((uint32_t)stride, (uint32_t)cap, (uint32_t)i, off) + m -
                                           kListFixed * kListLanes;
                if (R.staged) {
                    // An LDS index: back to the slot through the last stencil
                    // row that starts at or before it.
                    j = raw16[e];
                    int r = 8;
                    while (r > 0 && j < R.base[r]) --r;
                    j = (uint32_t)((int32_t)j - R.delta[r]);
                } else {
                    j = raw32[e];
                }
            }
            nbr[i * cap + m] = j;
        }
        off += list_tail_len(cnt[i]);
    }
    return FLUID_OK;
}

int fluid_sim_set_tile(FluidSim* s, uint32_t tile) {
    if (!valid(s)) return FLUID_ERR_ARG;
    const SolveVariant* v = solve_variant(tile);
    if (!v) {
        set_error("solve tile of %u slots: 64, 128 or 256", tile);
        return FLUID_ERR_ARG;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->V = v;
    s->lds_bytes = v->configure(s->solve, s->lds_limit);
    return FLUID_OK;
}

uint32_t fluid_sim_tile(const FluidSim* s) { return s && s->V ? s->V->tile : 0; }

int fluid_sim_set_tolerance(FluidSim* s, float tol, uint32_t min_passes, uint32_t max_passes, float div_tol,
                            uint32_t div_max_passes) {
    if (!valid(s)) return FLUID_ERR_ARG;
    if (tol > 0.0f && (max_passes == 0 || min_passes > max_passes || max_passes > kMaxPasses)) {
        set_error("tolerance mode needs 0 < min_passes <= max_passes <= %u", kMaxPasses);
        return FLUID_ERR_ARG;
    }
    if (div_tol > 0.0f && (div_max_passes == 0 || div_max_passes > kMaxPasses)) {
        set_error("divergence tolerance mode needs 0 < max_passes <= %u", kMaxPasses);
        return FLUID_ERR_ARG;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->tol.tol = tol > 0.0f ? tol : 0.0f;
    s->tol.min_passes = std::max<uint32_t>(min_passes, 1);
    s->tol.max_passes = max_passes;
    s->tol.div_tol = div_tol > 0.0f ? div_tol : 0.0f;
    s->tol.div_max = div_tol > 0.0f ? div_max_passes : 1u;
    return FLUID_OK;
}

int fluid_sim_iterations(FluidSim* s, uint32_t out[2]) {
    if (!valid(s) || !out) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(out, s->solve.iterations, 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(out + 1, s->solve.div_iterations, 4, hipMemcpyDeviceToHost));
    return FLUID_OK;
}

int fluid_sim_pass_histogram(FluidSim* s, uint32_t* out, bool reset) {
    if (!valid(s)) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (out) HIP_TRY(hipMemcpy(out, s->solve.pass_hist, 2 * kHistBins * 4, hipMemcpyDeviceToHost));
    if (reset) HIP_TRY(hipMemset(s->solve.pass_hist, 0, 2 * kHistBins * 4));
    return FLUID_OK;
}

int fluid_sim_track_ids(FluidSim* s, bool enable) {
    if (!valid(s)) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (!enable) {
        s->rest.ids = nullptr;
        s->work.ids = nullptr;
        s->track_ids = false;
        return FLUID_OK;
    }
    if (!s->track_ids) {
        uint32_t *a = nullptr, *b = nullptr;
        int rc = s->alloc(a, s->n);
        if (rc) return rc;
        if ((rc = s->alloc(b, s->n))) return rc;
        std::vector<uint32_t> iota(std::max<size_t>(s->n, 1));
        for (uint32_t i = 0; i < s->n; ++i) iota[i] = i;
        HIP_TRY(hipMemcpy(a, iota.data(), (size_t)s->n * 4, hipMemcpyHostToDevice));
        s->rest.ids = a;
        s->work.ids = b;
        s->track_ids = true;
    }
    return FLUID_OK;
}

int fluid_sim_download_ids(FluidSim* s, uint32_t* ids) {
    if (!valid(s) || !ids) return FLUID_ERR_ARG;
    if (!s->track_ids) {
        set_error("ids are not tracked");
        return FLUID_ERR_ARG;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(ids, s->rest.ids, (size_t)s->n * 4, hipMemcpyDeviceToHost));
    return FLUID_OK;
}

int fluid_sim_substeps(FluidSim* s, const uint8_t* step144, uint32_t n, uint32_t refine) {
    if (!valid(s) || !step144) return FLUID_ERR_ARG;
    return run_substeps(s, step144, n, refine);
}

int fluid_sim_density_only(FluidSim* s, const uint8_t* step144) {
    if (!valid(s) || !step144) return FLUID_ERR_ARG;
    StepBlock sb;
    std::memcpy(&sb, step144, sizeof sb);
    const OneScene sc{s->scene, make_step(sb, s->phys)};
    if (s->counted) {
        std::swap(s->grid.cellid, s->next_cellid);
        std::swap(s->grid.rank, s->next_rank);
    }
    launch_grid_build(s->stream, sc, s->n, s->cells, s->rest, s->work, s->grid, s->t(), s->counted);
    // The build consumed the count and no integrate follows to take the next.
    s->counted = false;
    s->V->density_only(s->stream, sc, s->n, s->grid.starts, s->work, s->solve, s->t());
    return check_launch("density_only");
}

int fluid_sim_reduce_stats(FluidSim* s) {
    if (!valid(s)) return FLUID_ERR_ARG;
    if (s->n == 0) return FLUID_OK;
    launch_reduce_stats(s->stream, s->params.rho0, s->n, s->solve.density, s->rest, s->solve.counters,
                        s->solve.clamps, s->stats,
                        s->t());
    return check_launch("reduce_stats");
}

int fluid_sim_read_stats(FluidSim* s, float* out14) {
    if (!valid(s) || !out14) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(out14, s->stats.stats, 14 * sizeof(float), hipMemcpyDeviceToHost));
    return FLUID_OK;
}

int fluid_sim_tracers(FluidSim* s, const uint8_t* step144) {
    if (!valid(s) || !step144) return FLUID_ERR_ARG;
    StepBlock sb;
    std::memcpy(&sb, step144, sizeof sb);
    launch_tracers(s->stream, s->scene, make_step(sb, s->phys), s->n, s->rest, s->tracers, s->t());
    return check_launch("tracers");
}

int fluid_sim_upload_tracers(FluidSim* s, const uint32_t* packed2) {
    if (!valid(s) || !packed2) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(s->tracers.tracers, packed2, (size_t)s->ntr * 8, hipMemcpyHostToDevice));
    return FLUID_OK;
}

int fluid_sim_download_tracers(FluidSim* s, uint32_t* packed2) {
    if (!valid(s) || !packed2) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(packed2, s->tracers.tracers, (size_t)s->ntr * 8, hipMemcpyDeviceToHost));
    return FLUID_OK;
}

int fluid_sim_frame(FluidSim* s, const uint8_t* step144, uint32_t n, uint32_t refine, const uint8_t* tracer_step144) {
    if (!valid(s) || !step144) return FLUID_ERR_ARG;
    StepBlock sb;
    std::memcpy(&sb, step144, sizeof sb);
    if (!(sb.dt > 0.0f)) return FLUID_OK;  // render.rs:3165-3167: no solve, no stats, no tracers
    int rc = run_substeps(s, step144, n, refine);
    if (rc) return rc;
    if ((rc = fluid_sim_reduce_stats(s))) return rc;
    if (s->ntr > 0 && tracer_step144) return fluid_sim_tracers(s, tracer_step144);
    return FLUID_OK;
}

int fluid_sim_counters(FluidSim* s, uint64_t* clamps, uint64_t* overflow) {
    if (!valid(s)) return FLUID_ERR_ARG;
    uint32_t c[4];
    std::vector<uint32_t> spread(kClampWords);
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(c, s->solve.counters, sizeof c, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(spread.data(), s->solve.clamps, kClampWords * 4, hipMemcpyDeviceToHost));
    uint64_t clamped = 0;
    for (uint32_t w = 0; w < kClampSlots; ++w) clamped += spread[w * kClampStride];
    if (clamps) *clamps = clamped;
    if (overflow) *overflow = c[1];
    return FLUID_OK;
}

int fluid_sim_params(const FluidSim* s, uint8_t* out64) {
    if (!valid(s) || !out64) return FLUID_ERR_ARG;
    std::memcpy(out64, &s->params, sizeof s->params);
    return FLUID_OK;
}

int fluid_sim_device_records(FluidSim* s, FluidDeviceRecords* out) {
    if (!valid(s) || !out) return FLUID_ERR_ARG;
    out->positions = reinterpret_cast<float*>(s->rest.pos);
    out->velocities = reinterpret_cast<float*>(s->rest.vel);
    out->prev_vel = reinterpret_cast<float*>(s->rest.prev_vel);
    out->prev_pressure = s->rest.prev_pressure;
    out->temperature = s->rest.temperature;
    // integrate writes the resting set at the working slot, so the last
    // substep's density is in the resting set's order.
    out->density = s->solve.density;
    out->tracers = s->ntr ? s->tracers.tracers : nullptr;
    out->count = s->n;
    out->tracer_count = s->ntr;
    out->stream = s->stream;
    return FLUID_OK;
}

int fluid_sim_set_physics(FluidSim* s, const float* phys5) {
    if (!valid(s) || !phys5) return FLUID_ERR_ARG;
    Physics ph;
    ph.mu = phys5[0];
    ph.k_near = phys5[1];
    ph.tension_scale = phys5[2];
    ph.adhesion_scale = phys5[3];
    ph.xsph_rate = phys5[4];
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->phys = ph;
    s->scene = make_scene(s->params, ph);
    return FLUID_OK;
}

int fluid_sim_synchronize(FluidSim* s) {
    if (!valid(s)) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FLUID_OK;
}

int fluid_sim_profile(FluidSim* s, uint32_t kernel_mask) { return fluid_sim_profile_sampled(s, kernel_mask, 1); }

int fluid_sim_profile_sampled(FluidSim* s, uint32_t kernel_mask, uint32_t stride) {
    if (!valid(s)) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->timer.reset();
    s->timer.mask = kernel_mask;
    s->timer.stride = stride ? stride : 1u;
    return FLUID_OK;
}

int fluid_sim_profile_read(FluidSim* s, double* ms, uint64_t* launches) {
    if (!valid(s)) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->timer.collect();
    for (int k = 0; k < KERNEL_COUNT_; ++k) {
        if (ms) ms[k] = s->timer.ms[k];
        if (launches) launches[k] = s->timer.launches[k];
    }
    return FLUID_OK;
}

int fluid_sim_describe(const FluidSim* s, char* buf, size_t size) {
    if (!valid(s) || !buf || size == 0) return FLUID_ERR_ARG;
    hipDeviceProp_t prop;
    const char* name = hipGetDeviceProperties(&prop, s->device) == hipSuccess ? prop.gcnArchName : "?";
    snprintf(buf, size,
             "{\"arch\": \"%s\", \"lds_bytes\": %zu, \"nbr_cap\": %u, \"stage_cap\": %u, "
             "\"list_lanes\": %u, \"tile\": %u}",
             name, s->lds_bytes, s->cap, s->solve.stage_cap, kListLanes, s->V->tile);
    return FLUID_OK;
}

const char* fluid_sim_kernel_name(uint32_t kernel) {
    return kernel < (uint32_t)KERNEL_COUNT_ ? kKernelNames[kernel] : "";
}

int fluid_scan_u32(const uint32_t* counts, uint32_t cells, uint32_t* starts) {
    if (!counts || !starts || cells == 0) {
        set_error("scan needs counts, starts and cells > 0");
        return FLUID_ERR_ARG;
    }
    if (fluid_sim_device_count() < 1) {
        set_error("no HIP device");
        return FLUID_ERR_DEVICE;
    }
    uint32_t *dc = nullptr, *ds = nullptr, *ticket = nullptr;
    uint64_t* status = nullptr;
    const uint32_t tiles = scan_tiles(cells);
    hipStream_t st = nullptr;
    int rc = FLUID_OK;
    hipError_t e = hipSuccess;
    do {
        if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) break;
        if ((e = hipMalloc(&dc, (size_t)cells * 4)) != hipSuccess) break;
        if ((e = hipMalloc(&ds, ((size_t)cells + 1) * 4)) != hipSuccess) break;
        if ((e = hipMalloc(&status, ((size_t)tiles + 1) * 8)) != hipSuccess) break;
        if ((e = hipMalloc(&ticket, 4)) != hipSuccess) break;
        if ((e = hipMemcpy(dc, counts, (size_t)cells * 4, hipMemcpyHostToDevice)) != hipSuccess) break;
        launch_scan_only(st, dc, ds, cells, status, ticket);
        if ((e = hipGetLastError()) != hipSuccess) break;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) break;
        if ((e = hipMemcpy(starts, ds, ((size_t)cells + 1) * 4, hipMemcpyDeviceToHost)) != hipSuccess) break;
    } while (false);
    if (e != hipSuccess) {
        set_error("scan failed: %s", hipGetErrorString(e));
        rc = FLUID_ERR_HIP;
    }
    (void)hipFree(dc);
    (void)hipFree(ds);
    (void)hipFree(status);
    (void)hipFree(ticket);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

}  // extern "C"
