// The stage-0 microbench behind fluid_renderer_create(bench_sweeps > 0):
// the reference's correctness-and-timing harness for the neighbour grid and
// one density sweep (crates/fluid-core/src/render.rs:2429-2729,
// sim_density.wgsl:36-99, sim_scan.wgsl:49-90).
//
// A seeded slab at the bench's spacing stays in seed order; each frame sorts
// it into an index list by cell (the grid build's count, look-back scan,
// place and stable rank, launch_index_sort in grid.hip, where the reference
// runs a three-pass scan limited to 65,536 cells and an atomic cursor), then
// runs `sweeps` density sweeps: each particle's SPH density from the 27
// neighbour cells through the index list, self included (sim_density.wgsl:
// 72-99). The first frame is checked on the host as validate_bench does
// (render.rs:2671-2729): the starts against a CPU counting scan of the seeded
// positions, every 16th density against a brute-force sum over all particles
// to 1e-3.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "fluid_common.hpp"
#include "fluid_kernels.hpp"
#include "host_common.hpp"

namespace fluid {

namespace {

// sim_density.wgsl:58-70: the cubic spline with sigma = 1 / (pi h^3) formed
// per call, r / h divided (the stage-0 shader's own form).
__device__ __forceinline__ float stage0_kernel(float r, float h) {
    const float q = r / h;
    const float sigma = 1.0f / (3.14159265f * h * h * h);
    if (q < 1.0f) return sigma * (1.0f - 1.5f * q * q * (1.0f - 0.5f * q));
    if (q < 2.0f) {
        const float t = 2.0f - q;
        return sigma * 0.25f * t * t * t;
    }
    return 0.0f;
}

// One thread a particle, in seed order: the 27 cells z, y, x outermost to
// innermost, each cell's particles through the index list.
__global__ __launch_bounds__(kBlock) void k_stage0_density(SceneK k, const float4* __restrict__ pos,
                                                           const uint32_t* __restrict__ starts,
                                                           const uint32_t* __restrict__ sorted,
                                                           float* __restrict__ density) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= k.count) return;
    const float4 p = pos[i];
    const int bx = (int)cell_coord(p.x, k.box_min[0], k.cell, k.dims[0]);
    const int by = (int)cell_coord(p.y, k.box_min[1], k.cell, k.dims[1]);
    const int bz = (int)cell_coord(p.z, k.box_min[2], k.cell, k.dims[2]);
    float rho = 0.0f;
    for (int dz = -1; dz <= 1; ++dz) {
        const int cz = bz + dz;
        if (cz < 0 || cz >= (int)k.dims[2]) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int cy = by + dy;
            if (cy < 0 || cy >= (int)k.dims[1]) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int cx = bx + dx;
                if (cx < 0 || cx >= (int)k.dims[0]) continue;
                const uint32_t c = ((uint32_t)cz * k.dims[1] + (uint32_t)cy) * k.dims[0] + (uint32_t)cx;
                const uint32_t end = starts[c + 1];
                for (uint32_t s = starts[c]; s < end; ++s) {
                    const float4 q = pos[sorted[s]];
                    const float ex = p.x - q.x, ey = p.y - q.y, ez = p.z - q.z;
                    const float r = sqrtf(ex * ex + ey * ey + ez * ez);
                    rho += k.mass * stage0_kernel(r, k.h);
                }
            }
        }
    }
    density[i] = rho;
}

}  // namespace

struct Stage0Bench {
    hipStream_t stream = nullptr;
    SimParamsBlock params{};
    SceneK scene{};
    FluidGrid grid{};
    std::vector<float> seeded;
    uint32_t n = 0, cells = 0, sweeps = 0;
    float4* pos = nullptr;
    float* density = nullptr;
    GridBufs g{};
    DeviceAllocs mem;
    int scan_ok = -1, density_ok = -1;
    double last_ms = 0.0;
    ~Stage0Bench() {
        if (stream) (void)hipStreamSynchronize(stream);
        mem.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// validate_bench (render.rs:2671-2729) on the host.
void validate(Stage0Bench* b) {
    std::vector<uint32_t> starts(b->cells + 1);
    std::vector<float> rho(b->n);
    if (hipMemcpy(starts.data(), b->g.starts, starts.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(rho.data(), b->density, rho.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        b->scan_ok = b->density_ok = 0;
        return;
    }
    std::vector<uint32_t> expect(b->cells, 0);
    for (uint32_t i = 0; i < b->n; ++i) expect[fluid_grid_cell_of(&b->grid, &b->seeded[4 * (size_t)i])] += 1;
    uint32_t run = 0;
    bool scan = true;
    for (uint32_t c = 0; c < b->cells; ++c) {
        scan = scan && starts[c] == run;
        run += expect[c];
    }
    b->scan_ok = scan ? 1 : 0;
    bool dens = true;
    const float* P = b->seeded.data();
    for (uint32_t i = 0; i < b->n && dens; i += 16) {
        float sum = 0.0f;
        for (uint32_t j = 0; j < b->n; ++j) {
            const float d0 = P[4 * i] - P[4 * j], d1 = P[4 * i + 1] - P[4 * j + 1], d2 = P[4 * i + 2] - P[4 * j + 2];
            sum += b->params.mass * fluid_kernel(std::sqrt(d0 * d0 + d1 * d1 + d2 * d2), b->params.h);
        }
        dens = std::fabs(rho[i] - sum) <= 1e-3f * sum;
    }
    b->density_ok = dens ? 1 : 0;
}

}  // namespace

Stage0Bench* stage0_create(const float extent[2], float spacing, uint32_t sweeps) {
    auto* b = new Stage0Bench();
    const float half[3] = {extent[0], extent[1], fluid_slab_depth() * 0.5f};
    const float h = 1.2f * spacing;
    fluid_grid_new(half, 2.0f * h, &b->grid);
    b->n = fluid_seed_slab(spacing, half, 0.5f, nullptr);
    b->seeded.resize((size_t)b->n * 4);
    fluid_seed_slab(spacing, half, 0.5f, b->seeded.data());
    const float rho0 = 998.2f;  // sim.rs:6
    uint8_t params[64];
    fluid_pack_sim_params(&b->grid, b->n, h, rho0 * spacing * spacing * spacing, rho0, params);
    std::memcpy(&b->params, params, sizeof b->params);
    b->scene = make_scene(b->params);
    b->cells = b->grid.dims[0] * b->grid.dims[1] * b->grid.dims[2];
    b->sweeps = sweeps;
    const size_t n = b->n, c = b->cells;
    int rc = FLUID_OK;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        (rc = b->mem.alloc(b->pos, n, false)) || (rc = b->mem.alloc(b->density, n)) ||
        (rc = b->mem.alloc(b->g.cellid, n)) || (rc = b->mem.alloc(b->g.rank, n)) ||
        (rc = b->mem.alloc(b->g.counts, c)) || (rc = b->mem.alloc(b->g.starts, c + 1)) ||
        (rc = b->mem.alloc(b->g.perm_tmp, n)) || (rc = b->mem.alloc(b->g.perm, n)) ||
        (rc = b->mem.alloc(b->g.status, scan_tiles(b->cells) + 1)) || (rc = b->mem.alloc(b->g.ticket, 1)) ||
        hipMemcpy(b->pos, b->seeded.data(), n * 16, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("stage-0 bench setup failed for %zu particles, %zu cells", n, c);
        delete b;
        return nullptr;
    }
    fprintf(stderr, "bench: %u particles, %u cells, %u sweeps, spacing %g m\n", b->n, b->cells, sweeps, spacing);
    return b;
}

void stage0_destroy(Stage0Bench* b) { delete b; }

int stage0_frame(Stage0Bench* b) {
    if (!b) return FLUID_ERR_ARG;
    const auto t0 = std::chrono::steady_clock::now();
    OneScene sc{b->scene, StepK{}};
    launch_index_sort(b->stream, sc, b->n, b->cells, b->pos, b->g);
    for (uint32_t s = 0; s < b->sweeps && b->n; ++s)
        hipLaunchKernelGGL(k_stage0_density, dim3(blocks_for(b->n)), dim3(kBlock), 0, b->stream, b->scene, b->pos,
                           b->g.starts, b->g.perm, b->density);
    int rc = check_launch("stage-0 frame");
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (b->scan_ok < 0) {
        validate(b);
        fprintf(stderr, "bench validation: scan %s, density %s\n", b->scan_ok ? "PASS" : "FAIL",
                b->density_ok ? "PASS" : "FAIL");
    }
    return FLUID_OK;
}

void stage0_validation(const Stage0Bench* b, int* scan_ok, int* density_ok) {
    if (scan_ok) *scan_ok = b ? b->scan_ok : -1;
    if (density_ok) *density_ok = b ? b->density_ok : -1;
}

uint32_t stage0_count(const Stage0Bench* b) { return b ? b->n : 0; }
uint32_t stage0_cells(const Stage0Bench* b) { return b ? b->cells : 0; }
double stage0_last_ms(const Stage0Bench* b) { return b ? b->last_ms : 0.0; }

int stage0_download(Stage0Bench* b, float* density, uint32_t* starts, float* pos4) {
    if (!b) return FLUID_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (density) HIP_TRY(hipMemcpy(density, b->density, (size_t)b->n * 4, hipMemcpyDeviceToHost));
    if (starts) HIP_TRY(hipMemcpy(starts, b->g.starts, ((size_t)b->cells + 1) * 4, hipMemcpyDeviceToHost));
    if (pos4) std::memcpy(pos4, b->seeded.data(), b->seeded.size() * 4);
    return FLUID_OK;
}

}  // namespace fluid
