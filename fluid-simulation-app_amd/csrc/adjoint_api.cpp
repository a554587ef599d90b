// Host side of the adjoint: a forward run of a fixed number of substeps that
// keeps a checkpoint of the resting set every k substeps, and a backward run
// that, segment by segment from the last, restores the checkpoint, replays
// the segment's substeps with a tape (the working set, the velocities and kd
// of each stage written to their own arrays by the kernels themselves, the
// neighbour lists, the permutation) and runs the reverse sweeps substep by
// substep.
//
// The taped replay runs with no LDS staging, so its lists hold slots; its
// results are bit-identical to the staged forward's (the staging only moves
// where the same values are read from). It runs on a second stream into one
// of two tape sets while the reverse sweeps read the other, so segment c - 1
// replays while segment c runs in reverse.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "adjoint.hpp"
#include "fluid_common.hpp"
#include "host_common.hpp"

using namespace fluid;

// Tape sets: the backward replays one segment into one set on the replay
// stream while the reverse sweeps read the other on the main stream.
constexpr int kTapeSets = 2;

struct FluidAdjoint {
    int device = 0;
    hipStream_t stream = nullptr;   // forward, reverse sweeps
    hipStream_t replay = nullptr;   // segment replays of the backward
    hipEvent_t ready[kTapeSets] = {}, done[kTapeSets] = {};
    SimParamsBlock params{};
    SceneK scene{};
    Physics phys{};
    uint32_t n = 0, cells = 0, cap = 0, steps = 0, every = 0, segments = 0;
    Records rest{}, work{};
    GridBufs grid{};
    SolveBufs solve{};
    std::vector<Records> checkpoints;
    std::vector<Tape> tapes[kTapeSets];
    std::vector<SolveBufs> tape_solve[kTapeSets];
    std::vector<Records> tape_work[kTapeSets];
    AdjointBufs L{};
    float* grad_g = nullptr;
    float* grad_mu = nullptr;
    std::vector<StepBlock> blocks;
    std::vector<uint32_t> refine;
    bool have_forward = false;
    hipEvent_t ev[4] = {};
    double forward_ms = 0.0, backward_ms = 0.0;
    DeviceAllocs mem;

    template <typename P>
    int alloc(P*& out, size_t count, bool zero = true) {
        return mem.alloc(out, count, zero);
    }
    ~FluidAdjoint() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (replay) (void)hipStreamSynchronize(replay);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        for (int i = 0; i < kTapeSets; ++i) {
            if (ready[i]) (void)hipEventDestroy(ready[i]);
            if (done[i]) (void)hipEventDestroy(done[i]);
        }
        mem.release();
        if (stream) (void)hipStreamDestroy(stream);
        if (replay) (void)hipStreamDestroy(replay);
    }
};

namespace {

bool valid(const FluidAdjoint* a) {
    if (!a) {
        set_error("null adjoint handle");
        return false;
    }
    return true;
}

int alloc_records(FluidAdjoint* A, Records& r, size_t n) {
    int rc;
    if ((rc = A->alloc(r.pos, n)) || (rc = A->alloc(r.vel, n)) || (rc = A->alloc(r.prev_vel, n)) ||
        (rc = A->alloc(r.prev_pressure, n)) || (rc = A->alloc(r.temperature, n)))
        return rc;
    r.ids = nullptr;
    return FLUID_OK;
}

int copy_records(hipStream_t st, const Records& from, const Records& to, size_t n) {
    HIP_TRY(hipMemcpyAsync(to.pos, from.pos, n * 16, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(to.vel, from.vel, n * 16, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(to.prev_vel, from.prev_vel, n * 16, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(to.prev_pressure, from.prev_pressure, n * 4, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(to.temperature, from.temperature, n * 4, hipMemcpyDeviceToDevice, st));
    return FLUID_OK;
}

int init_adjoint(FluidAdjoint* A, const uint8_t* params64, uint32_t steps, uint32_t every, uint32_t nbr_cap) {
    std::memcpy(&A->params, params64, sizeof A->params);
    const SimParamsBlock& p = A->params;
    if (p.dims[0] < 2 || p.dims[1] < 2 || p.dims[2] < 2 || p.count == 0 || steps == 0) {
        set_error("adjoint needs particles, a grid of at least 2 cells a side and at least one substep");
        return FLUID_ERR_ARG;
    }
    if (!(p.cell > 0.0f) || !(p.h > 0.0f)) {
        set_error("cell %g and h %g must be positive", p.cell, p.h);
        return FLUID_ERR_ARG;
    }
    A->cap = nbr_cap ? nbr_cap : FLUID_SIM_NBR_CAP;
    if (A->cap % kListLanes != 0 || A->cap > 256) {
        set_error("neighbour cap %u is not a multiple of 4 in [4, 256]", A->cap);
        return FLUID_ERR_ARG;
    }
    A->scene = make_scene(p, A->phys);
    A->n = p.count;
    A->cells = p.dims[0] * p.dims[1] * p.dims[2];
    A->steps = steps;
    A->every = every ? std::min(every, steps) : steps;
    A->segments = (steps + A->every - 1) / A->every;
    HIP_TRY(hipGetDevice(&A->device));
    HIP_TRY(hipStreamCreateWithFlags(&A->stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&A->replay, hipStreamNonBlocking));
    for (auto& e : A->ev) HIP_TRY(hipEventCreate(&e));
    for (int i = 0; i < kTapeSets; ++i) {
        HIP_TRY(hipEventCreateWithFlags(&A->ready[i], hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&A->done[i], hipEventDisableTiming));
    }
    const size_t n = A->n, c = A->cells;
    int rc = FLUID_OK;
#define A_(ptr, cnt)                           \
    if ((rc = A->alloc(ptr, cnt)) != FLUID_OK) \
        return rc;
    if ((rc = alloc_records(A, A->rest, n)) || (rc = alloc_records(A, A->work, n))) return rc;
    A_(A->grid.cellid, n);
    A_(A->grid.rank, n);
    A_(A->grid.counts, c);
    A_(A->grid.starts, c + 1);
    A_(A->grid.perm_tmp, n);
    A_(A->grid.perm, n);
    A_(A->grid.status, scan_tiles(A->cells) + 1);
    A_(A->grid.ticket, 1);
    SolveBufs& b = A->solve;
    A_(b.density, n);
    A_(b.alpha, n);
    A_(b.kd, n);
    A_(b.pressure, n);
    A_(b.rho_near, n);
    A_(b.accel, n);
    A_(b.xsph, n);
    A_(b.wall_grad, n);
    b.cap = A->cap;
    b.stride = std::max<uint32_t>(solve_tiles(A->n), 1) * kSolveTile;
    if ((rc = A->mem.alloc_list(b))) return rc;
    A_(b.nbr_n, n);
    A_(b.regions, std::max<uint32_t>(solve_tiles(A->n), 1));
    A_(b.counters, 4);
    A->grid.nonfinite = b.counters;
    A_(b.clamps, kClampWords);
    A_(b.zeros, kZeroRows);
    int lds = 0;
    HIP_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, A->device));
    configure_solve(b, (size_t)lds);
    A->checkpoints.resize(A->segments);
    for (auto& r : A->checkpoints)
        if ((rc = alloc_records(A, r, n))) return rc;
    // Tapes: each owns its working position and previous pressure, its
    // stage snapshots, its density-side fields and its lists.
    for (int set = 0; set < kTapeSets; ++set) {
    A->tapes[set].resize(A->every);
    A->tape_solve[set].resize(A->every);
    A->tape_work[set].resize(A->every);
    for (uint32_t t = 0; t < A->every; ++t) {
        Tape& T = A->tapes[set][t];
        A_(T.pos, n);
        A_(T.pp, n);
        for (auto& v : T.vel) A_(v, n);
        for (auto& k : T.kd) A_(k, n);
        A_(T.density, n);
        A_(T.rho_near, n);
        A_(T.alpha, n);
        A_(T.wall_grad, n);
        A_(T.accel, n);
        A_(T.perm, n);
        SolveBufs tb = b;
        tb.density = T.density;
        tb.rho_near = T.rho_near;
        tb.alpha = T.alpha;
        tb.wall_grad = T.wall_grad;
        tb.accel = T.accel;
        if ((rc = A->mem.alloc_list(tb))) return rc;
        A_(tb.nbr_n, n);
        tb.stage_cap = 0;
        A->tape_solve[set][t] = tb;
        Records w = A->work;
        w.pos = T.pos;
        w.vel = T.vel[0];
        w.prev_pressure = T.pp;
        A->tape_work[set][t] = w;
    }
    }
    AdjointBufs& L = A->L;
    A_(L.x, n);
    A_(L.v, n);
    A_(L.p, n);
    A_(L.x_out, n);
    A_(L.v_out, n);
    A_(L.pp_out, n);
    A_(L.kd, n);
    A_(L.D, n);
    A_(L.rho, n);
    A_(L.alpha, n);
    A_(L.pp, n);
    A_(L.a, n);
    A_(L.G, n);
    A_(L.xs, n);
    A_(L.w, n);
    A_(L.dfill, n);
    A_(L.J, 3 * n);
    A_(A->grad_g, 3 * (size_t)steps);
    A_(A->grad_mu, 1);
#undef A_
    return FLUID_OK;
}

int run_plain(FluidAdjoint* A, uint32_t s) {
    const OneScene sc{A->scene, make_step(A->blocks[s], A->phys)};
    launch_grid_build(A->stream, sc, A->n, A->cells, A->rest, A->work, A->grid, nullptr);
    launch_solve(A->stream, sc, A->n, A->grid.starts, A->work, A->rest, A->solve, A->refine[s], nullptr);
    return check_launch("adjoint forward");
}

// Substep s replayed on the replay stream into tape t of `set`.
int run_taped(FluidAdjoint* A, int set, uint32_t s, uint32_t t) {
    const OneScene sc{A->scene, make_step(A->blocks[s], A->phys)};
    GridBufs gb = A->grid;
    gb.perm = A->tapes[set][t].perm;
    const Records& wk = A->tape_work[set][t];
    launch_grid_build(A->replay, sc, A->n, A->cells, A->rest, wk, gb, nullptr);
    StageBufs stages;
    for (uint32_t i = 0; i < kMaxTapePasses + 3; ++i) stages.vel[i] = A->tapes[set][t].vel[i];
    for (uint32_t i = 0; i < kMaxTapePasses + 2; ++i) stages.kd[i] = A->tapes[set][t].kd[i];
    launch_solve_staged(A->replay, sc, A->n, A->grid.starts, wk, A->rest, A->tape_solve[set][t], A->refine[s],
                        stages);
    return check_launch("adjoint taped forward");
}

// Segment c replayed into tape set c % 2 once the reverse sweeps that last
// read that set are done; `ready` marks the set filled.
int replay_segment(FluidAdjoint* A, uint32_t c) {
    const int set = (int)(c % kTapeSets);
    const uint32_t s0 = c * A->every, s1 = std::min(s0 + A->every, A->steps);
    HIP_TRY(hipStreamWaitEvent(A->replay, A->done[set], 0));
    int rc;
    if ((rc = copy_records(A->replay, A->checkpoints[c], A->rest, A->n))) return rc;
    for (uint32_t s = s0; s < s1; ++s)
        if ((rc = run_taped(A, set, s, s - s0))) return rc;
    HIP_TRY(hipEventRecord(A->ready[set], A->replay));
    return FLUID_OK;
}

int upload_f4(hipStream_t st, float4* dst, const float* src, size_t n) {
    if (src) {
        HIP_TRY(hipMemcpyAsync(dst, src, n * 16, hipMemcpyHostToDevice, st));
    } else {
        adj_zero(st, reinterpret_cast<float*>(dst), n * 4);
    }
    return FLUID_OK;
}

}  // namespace

extern "C" {

FluidAdjoint* fluid_adjoint_create(const uint8_t* params64, uint32_t steps, uint32_t checkpoint_every,
                                   uint32_t nbr_cap) {
    if (!params64) {
        set_error("null SimParams block");
        return nullptr;
    }
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices < 1) {
        set_error("no HIP device");
        return nullptr;
    }
    FluidAdjoint* A = new FluidAdjoint();
    if (init_adjoint(A, params64, steps, checkpoint_every, nbr_cap) != FLUID_OK) {
        delete A;
        return nullptr;
    }
    return A;
}

void fluid_adjoint_destroy(FluidAdjoint* a) { delete a; }

int fluid_adjoint_set_physics(FluidAdjoint* A, const float* phys5) {
    if (!valid(A) || !phys5) return FLUID_ERR_ARG;
    Physics ph;
    ph.mu = phys5[0];
    ph.k_near = phys5[1];
    ph.tension_scale = phys5[2];
    ph.adhesion_scale = phys5[3];
    ph.xsph_rate = phys5[4];
    A->phys = ph;
    A->scene = make_scene(A->params, ph);
    A->have_forward = false;
    return FLUID_OK;
}

int fluid_adjoint_forward(FluidAdjoint* A, const float* pos4, const float* vel4, const uint8_t* steps144,
                          const uint32_t* refine, float* pos_out4, float* vel_out4) {
    if (!valid(A) || !pos4 || !steps144 || !refine) return FLUID_ERR_ARG;
    A->blocks.resize(A->steps);
    A->refine.assign(refine, refine + A->steps);
    for (uint32_t s = 0; s < A->steps; ++s) {
        std::memcpy(&A->blocks[s], steps144 + 144 * (size_t)s, sizeof(StepBlock));
        if (A->refine[s] > kMaxTapePasses || !(A->blocks[s].dt > 0.0f)) {
            set_error("substep %u: the adjoint takes dt > 0 and at most %u density passes", s, kMaxTapePasses);
            return FLUID_ERR_ARG;
        }
        if (A->blocks[s].touch_n != 0) {
            set_error("substep %u has %u touches: the adjoint has no reverse of the finger pull", s,
                      A->blocks[s].touch_n);
            return FLUID_ERR_ARG;
        }
    }
    const size_t n = A->n;
    HIP_TRY(hipStreamSynchronize(A->stream));
    HIP_TRY(hipMemcpy(A->rest.pos, pos4, n * 16, hipMemcpyHostToDevice));
    int rc;
    if ((rc = upload_f4(A->stream, A->rest.vel, vel4, n))) return rc;
    adj_zero(A->stream, reinterpret_cast<float*>(A->rest.prev_vel), n * 4);
    adj_zero(A->stream, A->rest.prev_pressure, n);
    std::vector<float> warm(n, 293.15f);
    HIP_TRY(hipMemcpyAsync(A->rest.temperature, warm.data(), n * 4, hipMemcpyHostToDevice, A->stream));
    adj_zero(A->stream, reinterpret_cast<float*>(A->solve.counters), 4);
    HIP_TRY(hipEventRecord(A->ev[0], A->stream));
    for (uint32_t s = 0; s < A->steps; ++s) {
        if (s % A->every == 0 && (rc = copy_records(A->stream, A->rest, A->checkpoints[s / A->every], n)))
            return rc;
        if ((rc = run_plain(A, s))) return rc;
    }
    HIP_TRY(hipEventRecord(A->ev[1], A->stream));
    HIP_TRY(hipStreamSynchronize(A->stream));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, A->ev[0], A->ev[1]));
    A->forward_ms = ms;
    uint32_t counters[4];
    HIP_TRY(hipMemcpy(counters, A->solve.counters, sizeof counters, hipMemcpyDeviceToHost));
    if (counters[1]) {
        set_error("a neighbour list overflowed its cap (%u pairs): the adjoint needs whole, symmetric lists",
                  counters[1]);
        return FLUID_ERR_ARG;
    }
    if (pos_out4) HIP_TRY(hipMemcpy(pos_out4, A->rest.pos, n * 16, hipMemcpyDeviceToHost));
    if (vel_out4) HIP_TRY(hipMemcpy(vel_out4, A->rest.vel, n * 16, hipMemcpyDeviceToHost));
    A->have_forward = true;
    return FLUID_OK;
}

int fluid_adjoint_backward(FluidAdjoint* A, const float* dpos4, const float* dvel4, float* dvel0_4,
                           float* dpos0_4, float* dgravity3, float* dmu) {
    if (!valid(A) || !dpos4) return FLUID_ERR_ARG;
    if (!A->have_forward) {
        set_error("fluid_adjoint_backward needs a forward run first");
        return FLUID_ERR_ARG;
    }
    const size_t n = A->n;
    AdjointBufs L = A->L;
    HIP_TRY(hipStreamSynchronize(A->stream));
    HIP_TRY(hipMemcpyAsync(L.x, dpos4, n * 16, hipMemcpyHostToDevice, A->stream));
    int rc;
    if ((rc = upload_f4(A->stream, L.v, dvel4, n))) return rc;
    adj_zero(A->stream, L.p, n);
    adj_zero(A->stream, A->grad_g, 3 * (size_t)A->steps);
    adj_zero(A->stream, A->grad_mu, 1);
    L.mu = A->grad_mu;
    HIP_TRY(hipEventRecord(A->ev[2], A->stream));
    // The replay stream starts after this call's uploads; the tape sets are
    // free (nothing reads them yet).
    HIP_TRY(hipStreamWaitEvent(A->replay, A->ev[2], 0));
    for (int set = 0; set < kTapeSets; ++set) HIP_TRY(hipEventRecord(A->done[set], A->stream));
    if ((rc = replay_segment(A, A->segments - 1))) return rc;
    for (uint32_t c = A->segments; c-- > 0;) {
        // the next segment back replays while this one runs in reverse
        if (c > 0 && (rc = replay_segment(A, c - 1))) return rc;
        const int set = (int)(c % kTapeSets);
        const uint32_t s0 = c * A->every, s1 = std::min(s0 + A->every, A->steps);
        HIP_TRY(hipStreamWaitEvent(A->stream, A->ready[set], 0));
        for (uint32_t s = s1; s-- > s0;) {
            const uint32_t t = s - s0;
            L.g = A->grad_g + 3 * (size_t)s;
            launch_reverse_substep(A->stream, A->scene, make_step(A->blocks[s], A->phys), A->n,
                                   A->refine[s], A->tapes[set][t], A->tape_solve[set][t], L);
            std::swap(L.x, L.x_out);
            std::swap(L.v, L.v_out);
            std::swap(L.p, L.pp_out);
            if ((rc = check_launch("adjoint reverse"))) return rc;
        }
        HIP_TRY(hipEventRecord(A->done[set], A->stream));
    }
    HIP_TRY(hipEventRecord(A->ev[3], A->stream));
    HIP_TRY(hipStreamSynchronize(A->stream));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, A->ev[2], A->ev[3]));
    A->backward_ms = ms;
    if (dvel0_4) HIP_TRY(hipMemcpy(dvel0_4, L.v, n * 16, hipMemcpyDeviceToHost));
    if (dpos0_4) HIP_TRY(hipMemcpy(dpos0_4, L.x, n * 16, hipMemcpyDeviceToHost));
    if (dgravity3) HIP_TRY(hipMemcpy(dgravity3, A->grad_g, 3 * (size_t)A->steps * 4, hipMemcpyDeviceToHost));
    if (dmu) HIP_TRY(hipMemcpy(dmu, A->grad_mu, 4, hipMemcpyDeviceToHost));
    return FLUID_OK;
}

int fluid_adjoint_timing(const FluidAdjoint* A, double* forward_ms, double* backward_ms) {
    if (!valid(A)) return FLUID_ERR_ARG;
    if (forward_ms) *forward_ms = A->forward_ms;
    if (backward_ms) *backward_ms = A->backward_ms;
    return FLUID_OK;
}

}  // extern "C"
