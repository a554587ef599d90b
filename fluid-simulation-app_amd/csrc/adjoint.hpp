// Internal interface of the adjoint: what the taped forward records for one
// substep, and the adjoint buffers the reverse sweeps update.
#pragma once

#include "fluid_kernels.hpp"

namespace fluid {

// The most density passes a taped substep records.
constexpr uint32_t kMaxTapePasses = kMaxStagePasses;

// One substep of the taped forward, all in that substep's working (slot)
// order. vel[0] enters div_apply, vel[1] leaves it, vel[2] leaves
// forces_den_apply, vel[3 + r] leaves den_apply of pass r; kd[0] is the
// divergence kappa, kd[1] the warm start, kd[2 + r] den_kappa of pass r.
// perm maps each slot to the previous resting index. The neighbour lists are
// in the substep's own SolveBufs (slots, not LDS indices).
struct Tape {
    float4* pos;
    float* pp;
    float4* vel[kMaxTapePasses + 3];
    float* kd[kMaxTapePasses + 2];
    float* density;
    float* rho_near;
    float* alpha;
    float4* wall_grad;
    float4* accel;
    uint32_t* perm;
};

// Adjoint (gradient) buffers, working order unless noted.
struct AdjointBufs {
    float4 *x, *v;            // of the resting set integrate wrote, then of the working position / velocity
    float* p;                 // of the pressure integrate carried over
    float4 *x_out, *v_out;    // of the previous resting set (its order)
    float* pp_out;
    float *kd, *D, *rho, *alpha, *pp;
    float4 *a, *G;            // G.w: the adjoint of gsq
    float4* xs;               // of the XSPH blend forces_eval leaves
    float* w;                 // of rho_near (the near-pressure density)
    float4 *dfill, *J;        // grad fill; d wall_grad / d x_a (3 a particle)
    float* g;                 // [3] this substep's gravity adjoint (added to)
    float* mu;                // [1] the viscosity adjoint (added to)
};

void adj_zero(hipStream_t st, float* a, size_t count);
void launch_reverse_substep(hipStream_t st, const SceneK& k, const StepK& s, uint32_t n, uint32_t passes,
                            const Tape& T, const SolveBufs& b, const AdjointBufs& L);

}  // namespace fluid
