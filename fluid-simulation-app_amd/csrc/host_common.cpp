// Host-side helpers shared by the single-scene and batched C ABI.
#include "host_common.hpp"

namespace fluid {

namespace {
thread_local std::string g_error;
}

void set_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    fprintf(stderr, "fluid_sim: %s\n", buf);
}

const char* last_error() { return g_error.c_str(); }

const char* kKernelNames[KERNEL_COUNT_] = {
    "grid_count", "grid_scan", "grid_scatter", "density_div", "div_apply", "forces",
    "unused", "den_kappa", "den_apply", "integrate", "reduce_stats", "tracers"};

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: kernel launch failed: %s", what, hipGetErrorString(e));
        return FLUID_ERR_HIP;
    }
    return FLUID_OK;
}

const SolveVariant* solve_variant(uint32_t tile) {
    switch (tile) {
        case 64: return &t64::variant();
        case 128: return &t128::variant();
        case 256: return &t256::variant();
        default: return nullptr;
    }
}

}  // namespace fluid
