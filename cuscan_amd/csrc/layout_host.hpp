// Host-side helpers of the layout driver shared by the plan, the batched
// layouts and the transform server.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "stdrng.hpp"

namespace umaphip {

// adam_schedule / momentum_schedule (umap-core/src/modern_layout.rs:110-165),
// converted once to f32 (Schedule64::to_f32, :98-107).
void schedules(int epochs, double alpha0, bool good, bool momentum, std::vector<float> &alpha,
               std::vector<float> &mu);
// Adam bias corrections c1 = f32(1 - (0.9f32 as f64)^(t+1)), c2 likewise
// with 0.99f32 (modern_runtime.rs:1286-1287).
void bias_corrections(int epochs, std::vector<float> &c1, std::vector<float> &c2);
// effective_negative_range (umap-core/src/modern_layout.rs:450-457).
uint64_t effective_range(uint64_t raw, uint64_t refs);
// resolved_negative_scale (umap-core/src/modern_layout.rs:537-561).
float resolved_scale(uint64_t raw, uint64_t refs, bool has_explicit, float explicit_scale,
                     bool adaptation);
// ProjectionOrder direction (modern_negatives.rs:208-218, 283-297).
void projection_direction(StdRng &rng, int64_t d, float *dir);

// Host replay of ModernOrders (modern_orders.rs:359-680) for a layout.
struct OrderReplay {
    StdRng rng;
    std::vector<int32_t> source, negative;
    bool alias, negative_live, source_needed;
    OrderReplay(int64_t h, int64_t n, uint64_t seed, bool alias_, bool hard, bool momentum)
        : rng(seed), alias(alias_) {
        // The negative order is shuffled only while neither aliased nor retired
        // (retirement happens before the first shuffle in projection mode).
        negative_live = !alias && !hard;
        // Adam never reads the source order (whole-epoch Jacobi), so its
        // shuffles only matter for the words they consume before a live
        // negative shuffle.
        source_needed = momentum;
        source.resize(static_cast<size_t>(h));
        for (int64_t i = 0; i < h; ++i) source[i] = static_cast<int32_t>(i);
        if (!alias) {
            negative.resize(static_cast<size_t>(n));
            for (int64_t i = 0; i < n; ++i) negative[i] = static_cast<int32_t>(i);
        }
    }
    bool active() const { return source_needed || negative_live; }
    // prepare_next: source first, then the live negative order.
    void advance() {
        if (!active()) return;
        rng.shuffle(source.data(), source.size(), source_needed);
        if (negative_live) rng.shuffle(negative.data(), negative.size(), true);
    }
    const std::vector<int32_t> &neg() const { return alias ? source : negative; }
};

// Pinned (page-locked) host staging buffer.
struct Pinned {
    void *ptr = nullptr;
    explicit Pinned(size_t bytes) {
        if (bytes) HIP_CHECK(hipHostMalloc(&ptr, bytes, hipHostMallocDefault));
    }
    ~Pinned() {
        if (ptr) (void)hipHostFree(ptr);
    }
    Pinned(const Pinned &) = delete;
    Pinned &operator=(const Pinned &) = delete;
};

}  // namespace umaphip
