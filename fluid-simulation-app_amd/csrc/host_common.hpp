// Host-side helpers shared by the single-scene and batched C ABI: the
// error channel, HIP call checking, device allocation and the per-kernel
// HIP-event timer.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "fluid_kernels.hpp"
#include "../../include/fluid_sim.h"

namespace fluid {

// The last error message of this thread (fluid_sim_last_error); also
// printed to stderr.
void set_error(const char* fmt, ...);
const char* last_error();

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::fluid::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return FLUID_ERR_HIP;                                                       \
        }                                                                               \
    } while (0)

extern const char* kKernelNames[KERNEL_COUNT_];

// FLUID_ERR_HIP with the message if the last launch failed.
int check_launch(const char* what);

// Device buffers of one handle, freed together.
struct DeviceAllocs {
    std::vector<void*> ptrs;
    template <typename P>
    int alloc(P*& out, size_t count, bool zero = true) {
        void* p = nullptr;
        size_t bytes = std::max<size_t>(count, 1) * sizeof(P);
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
            return FLUID_ERR_NOMEM;
        }
        ptrs.push_back(p);
        if (zero) {
            e = hipMemset(p, 0, bytes);
            if (e != hipSuccess) {
                set_error("hipMemset failed: %s", hipGetErrorString(e));
                return FLUID_ERR_HIP;
            }
        }
        out = static_cast<P*>(p);
        return FLUID_OK;
    }
    // The neighbour list of `cap` entries for each of `stride` slots.
    int alloc_list(SolveBufs& b) {
        const size_t e = (size_t)b.cap * b.stride;
        // The kernels index the list with 32-bit offsets.
        if (e + kListPad >= ((size_t)1 << 32)) {
            set_error("neighbour list of %zu entries (cap %u x %u slots) exceeds 2^32", e, b.cap, b.stride);
            return FLUID_ERR_ARG;
        }
        // ... and the per-slot float4 arrays with 32-bit byte offsets.
        if ((size_t)b.stride * sizeof(float4) >= ((size_t)1 << 32)) {
            set_error("%u slots exceed the 2^28 a sweep addresses", b.stride);
            return FLUID_ERR_ARG;
        }
        int rc;
        const size_t ep = e + kListPad;
        if ((rc = alloc(b.nbr16, ep, false)) || (rc = alloc(b.nbr32, ep, false)) || (rc = alloc(b.nbrg, ep, false)))
            return rc;
        // One spill area serves every list that copies these buffers.
        const size_t tail = (size_t)list_tail_cap(b.cap) * b.stride;
        if (!b.spill && tail && (rc = alloc(b.spill, tail, false))) return rc;
        return FLUID_OK;
    }
    void release() {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
    }
};

class EventTimer final : public KernelTimer {
  public:
    uint32_t mask = 0;
    // Times one launch in every `stride` of each class: a timed launch's
    // events go into the queue beside it and slow the stream a little.
    uint32_t stride = 1;
    uint64_t seen[KERNEL_COUNT_] = {};
    struct Pair {
        hipEvent_t a, b;
        int kernel;
    };
    std::vector<Pair> pending;
    std::vector<hipEvent_t> pool;
    hipEvent_t open[KERNEL_COUNT_] = {};
    double ms[KERNEL_COUNT_] = {};
    uint64_t launches[KERNEL_COUNT_] = {};

    ~EventTimer() override {
        for (auto& p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
    bool wants(int kernel) const { return (mask >> kernel) & 1u; }
    // Whether this launch of the class is one to time.
    bool sample(int kernel) { return seen[kernel]++ % (stride ? stride : 1u) == 0u; }
    void begin(hipStream_t st, int kernel) override {
        if (!wants(kernel) || !sample(kernel)) return;
        open[kernel] = take();
        (void)hipEventRecord(open[kernel], st);
    }
    void end(hipStream_t st, int kernel) override {
        if (!wants(kernel) || !open[kernel]) return;
        hipEvent_t e = take();
        (void)hipEventRecord(e, st);
        pending.push_back({open[kernel], e, kernel});
        open[kernel] = nullptr;
    }
    hipEvent_t start(int kernel) override {
        if (!wants(kernel) || !sample(kernel)) return nullptr;
        open[kernel] = take();
        return open[kernel];
    }
    hipEvent_t stop(int kernel) override {
        if (!wants(kernel) || !open[kernel]) return nullptr;
        hipEvent_t e = take();
        pending.push_back({open[kernel], e, kernel});
        open[kernel] = nullptr;
        return e;
    }
    void collect() {
        for (auto& p : pending) {
            (void)hipEventSynchronize(p.b);
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, p.a, p.b);
            ms[p.kernel] += t;
            launches[p.kernel] += 1;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    void reset() {
        collect();
        std::fill(ms, ms + KERNEL_COUNT_, 0.0);
        std::fill(launches, launches + KERNEL_COUNT_, 0);
        std::fill(seen, seen + KERNEL_COUNT_, 0);
    }
};


}  // namespace fluid
