// Compatibility (classic) SGD layout: cuscan's optimizer="compatibility".
//
// Semantics follow crates/umap/src/kernels/standard_umap/layout.rs (the
// atomic vertex kernel, :68-336) and its driver
// crates/umap/src/standard_umap/estimator.rs (optimize :2868-2982,
// prepare_schedule :2753-2798, optimizer_rng_from_first_coordinates
// :2800-2829, build_head_csr :3080-3131). Per epoch:
//   1. one lane per edge advances that edge's f64 clocks and records a fire
//      mask (-1: no positive event, else the number of negatives, >= 0);
//   2. one lane per head vertex walks the vertex's positive edges in edge-id
//      order: an attraction moves head and (fit) tail, each negative draws
//      from the vertex's 3-word Tausworthe generator and moves the head.
//      Coordinates are read live and updated with atomic adds (Hogwild), so
//      the result is not unique; the contract is embedding quality. Every
//      vertex consumes exactly its serial draw sequence, so the generator
//      state after each epoch and the clocks are deterministic.
// Fit runs the vertex pass as four ascending contiguous head batches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "stdrng.hpp"

namespace umaphip {
namespace {

struct ClassicArgs {
    int64_t h, n, d, E;
    float *pos;             // [h,d] head, updated in place
    const float *tail;      // [n,d] fixed tail (transform); pos when shared
    const int32_t *tails;   // [E]
    const float *weights;   // [E]
    const double *per;      // [E] epochs per sample
    double *npos;           // [E]
    const double *nper;     // [E] epochs per negative sample
    double *nneg;           // [E]
    int32_t *mask;          // [E] scratch
    int64_t *rng;           // [h,3]
    const int32_t *head_offsets, *head_edges;
    uint32_t epoch;
    float a, b, gamma, alpha, rate;
    int shared;
};

__global__ void classic_clock_kernel(ClassicArgs A) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= A.E) return;
    int32_t fired = -1;
    const double now = static_cast<double>(A.epoch);
    if (A.weights[e] > 0.0f && A.npos[e] <= now) {
        fired = 0;
        A.npos[e] = A.npos[e] + A.per[e];
        if (A.rate > 0.0f) {
            const double clock = A.nneg[e], interval = A.nper[e];
            const int32_t count = static_cast<int32_t>((now - clock) / interval);
            if (count > 0) fired = count;
            A.nneg[e] = clock + static_cast<double>(count) * interval;
        }
    }
    A.mask[e] = fired;
}

// One Tausworthe component on the signed 64-bit state: the masked left part
// is confined to 32 bits, the right part keeps the state's high bits and
// shifts arithmetically.
__device__ __host__ inline int64_t taus_step(int64_t s, uint64_t keep, int l, int m, int r) {
    const uint64_t u = static_cast<uint64_t>(s);
    const uint64_t left = ((u & keep) << l) & 0xFFFFFFFFull;
    const int64_t right = static_cast<int64_t>(((u << m) & 0xFFFFFFFFull) ^ u) >> r;
    return static_cast<int64_t>(left) ^ right;
}

__device__ inline float load_coord(const float *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline float classic_distance(const ClassicArgs &A, int64_t i, int64_t j) {
    float q = 0.0f;
    for (int64_t c = 0; c < A.d; ++c) {
        const float x = load_coord(A.pos + i * A.d + c);
        const float y = A.shared ? load_coord(A.pos + j * A.d + c) : A.tail[j * A.d + c];
        const float diff = x - y;
        q = q + diff * diff;
    }
    return q;
}

__device__ inline void classic_apply(const ClassicArgs &A, int64_t i, int64_t j, float coef,
                                     bool move_tail) {
    for (int64_t c = 0; c < A.d; ++c) {
        const float x = load_coord(A.pos + i * A.d + c);
        const float y = A.shared ? load_coord(A.pos + j * A.d + c) : A.tail[j * A.d + c];
        float g = coef * (x - y);
        g = g < -4.0f ? -4.0f : (g > 4.0f ? 4.0f : g);
        const float update = g * A.alpha;
        atomicAdd(A.pos + i * A.d + c, update);
        if (A.shared && move_tail) atomicAdd(A.pos + j * A.d + c, -update);
    }
}

__global__ void classic_vertex_kernel(ClassicArgs A, int64_t start, int64_t end) {
    const int64_t v = start + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= end) return;
    const int64_t n_tail = A.n;
    int64_t r0 = A.rng[v * 3], r1 = A.rng[v * 3 + 1], r2 = A.rng[v * 3 + 2];
    bool used = false;
    for (int32_t slot = A.head_offsets[v]; slot < A.head_offsets[v + 1]; ++slot) {
        const int32_t edge = A.head_edges[slot];
        const int32_t fired = A.mask[edge];
        if (fired < 0) continue;
        const int64_t t = static_cast<uint32_t>(A.tails[edge]);
        if (t < n_tail && (!A.shared || t != v)) {
            const float q = classic_distance(A, v, t);
            if (q > 0.0f) {
                const float coef = (-2.0f * A.a * A.b) * powf(q, A.b - 1.0f) /
                                   (A.a * powf(q, A.b) + 1.0f);
                classic_apply(A, v, t, coef, true);
            }
        }
        for (int32_t k = 0; k < fired; ++k) {
            used = true;
            r0 = taus_step(r0, 4294967294ull, 12, 13, 19);
            r1 = taus_step(r1, 4294967288ull, 4, 2, 25);
            r2 = taus_step(r2, 4294967280ull, 17, 3, 11);
            const int32_t draw = static_cast<int32_t>(static_cast<uint32_t>(
                static_cast<uint64_t>(r0 ^ r1 ^ r2)));
            int32_t rem = draw % static_cast<int32_t>(n_tail);
            if (rem < 0) rem += static_cast<int32_t>(n_tail);
            const int64_t s = rem;
            if (A.shared && s == v) continue;
            const float q = classic_distance(A, v, s);
            if (q > 0.0f) {
                const float coef = (2.0f * A.gamma * A.b) /
                                   ((0.001f + q) * (A.a * powf(q, A.b) + 1.0f));
                classic_apply(A, v, s, coef, false);
            }
        }
    }
    if (used) {
        A.rng[v * 3] = r0;
        A.rng[v * 3 + 1] = r1;
        A.rng[v * 3 + 2] = r2;
    }
}

void classic_epoch_launch(const ClassicArgs &A, hipStream_t s) {
    if (A.E > 0) {
        hipLaunchKernelGGL(classic_clock_kernel, dim3(static_cast<unsigned>((A.E + 255) / 256)),
                           dim3(256), 0, s, A);
        HIP_CHECK(hipGetLastError());
    }
    auto batch = [&](int64_t start, int64_t end) {
        if (start >= end) return;
        hipLaunchKernelGGL(classic_vertex_kernel,
                           dim3(static_cast<unsigned>((end - start + 255) / 256)), dim3(256), 0, s,
                           A, start, end);
        HIP_CHECK(hipGetLastError());
    };
    if (A.shared) {
        const int64_t size = (A.h + 3) / 4;
        for (int k = 0; k < 4; ++k) batch(k * size, std::min(A.h, (k + 1) * size));
    } else {
        batch(0, A.h);
    }
}

}  // namespace
}  // namespace umaphip

using namespace umaphip;

extern "C" {

umaphip_status umaphip_classic_epoch(const umaphip_classic_args *args, void *stream) {
    return guarded([&] {
        UMAPHIP_REQUIRE(args != nullptr, UMAPHIP_INVALID_INPUT, "classic args are null");
        const umaphip_classic_args &a = *args;
        UMAPHIP_REQUIRE(a.h >= 0 && a.n > 0 && a.d > 0 && a.n_edges >= 0, UMAPHIP_INVALID_INPUT,
                        "classic layout sizes must be positive");
        UMAPHIP_REQUIRE(a.h <= INT32_MAX && a.n <= INT32_MAX && a.n_edges <= INT32_MAX,
                        UMAPHIP_OVERFLOW, "classic layout sizes must fit i32");
        UMAPHIP_REQUIRE(a.head && a.head_offsets && a.rng_state &&
                            (a.n_edges == 0 || (a.tails && a.weights && a.epochs_per_sample &&
                                                a.epoch_of_next_sample &&
                                                a.epochs_per_negative_sample &&
                                                a.epoch_of_next_negative_sample && a.mask &&
                                                a.head_edges)),
                        UMAPHIP_INVALID_INPUT, "classic layout arrays are null");
        UMAPHIP_REQUIRE(a.shared || a.tail, UMAPHIP_INVALID_INPUT, "transform requires a tail");
        UMAPHIP_REQUIRE(!a.shared || a.n == a.h, UMAPHIP_INVALID_INPUT,
                        "shared layouts have as many tail rows as head rows");
        UMAPHIP_REQUIRE((a.move_other != 0) == (a.shared != 0), UMAPHIP_INVALID_CONFIG,
                        "the layout supports shared fit (move_other) or fixed-tail transform only");
        ClassicArgs A{};
        A.h = a.h; A.n = a.n; A.d = a.d; A.E = a.n_edges;
        A.pos = a.head; A.tail = a.shared ? a.head : a.tail;
        A.tails = a.tails; A.weights = a.weights;
        A.per = a.epochs_per_sample; A.npos = a.epoch_of_next_sample;
        A.nper = a.epochs_per_negative_sample; A.nneg = a.epoch_of_next_negative_sample;
        A.mask = a.mask; A.rng = a.rng_state;
        A.head_offsets = a.head_offsets; A.head_edges = a.head_edges;
        A.epoch = a.epoch; A.a = a.a; A.b = a.b; A.gamma = a.gamma; A.alpha = a.alpha;
        A.rate = a.negative_sample_rate; A.shared = a.shared != 0;
        classic_epoch_launch(A, static_cast<hipStream_t>(stream));
    });
}

umaphip_status umaphip_classic_optimize(int64_t h, int64_t n, int64_t d, float *head,
                                        const float *tail, const int32_t *heads,
                                        const int32_t *tails, const float *weights, int64_t E,
                                        int32_t n_epochs, int32_t default_epochs, float a,
                                        float b, float learning_rate, float gamma,
                                        int32_t negative_sample_rate, uint64_t seed,
                                        int32_t shared, double *next_positive_out,
                                        double *next_negative_out, int64_t *rng_out) {
    return guarded([&] {
        if (shared) n = h;
        UMAPHIP_REQUIRE(h > 0 && n > 0 && d > 0 && E >= 0 && n_epochs >= 0 && default_epochs > 0,
                        UMAPHIP_INVALID_INPUT, "classic layout sizes must be positive");
        UMAPHIP_REQUIRE(h <= INT32_MAX && n <= INT32_MAX && E <= INT32_MAX, UMAPHIP_OVERFLOW,
                        "classic layout sizes must fit i32");
        UMAPHIP_REQUIRE(head && (shared || tail) && (E == 0 || (heads && tails && weights)),
                        UMAPHIP_INVALID_INPUT, "classic layout arrays are null");
        UMAPHIP_REQUIRE(negative_sample_rate >= 0, UMAPHIP_INVALID_CONFIG,
                        "negative sample rate must be nonnegative");
        // prepare_schedule: prune below max / divisor, f64 intervals max / w
        float wmax = 0.0f;
        for (int64_t e = 0; e < E; ++e) {
            UMAPHIP_REQUIRE(std::isfinite(weights[e]) && weights[e] >= 0.0f, UMAPHIP_NUMERICAL,
                            "layout edge has invalid weight");
            wmax = std::max(wmax, weights[e]);
        }
        UMAPHIP_REQUIRE(E == 0 || wmax > 0.0f, UMAPHIP_INVALID_INPUT,
                        "fuzzy graph is empty after positive-edge pruning");
        const float divisor = static_cast<float>(n_epochs > 10 ? n_epochs : default_epochs);
        const float threshold = wmax / divisor;
        std::vector<float> pruned(static_cast<size_t>(E));
        std::vector<double> per(E), npos(E), nper(E), nneg(E);
        const double rate = static_cast<double>(std::max<int32_t>(negative_sample_rate, 1));
        std::vector<int32_t> counts(static_cast<size_t>(h), 0);
        for (int64_t e = 0; e < E; ++e) {
            const float w = weights[e] < threshold ? 0.0f : weights[e];
            pruned[e] = w;
            const double safe = w == 0.0f ? 1.0 : static_cast<double>(w);
            per[e] = static_cast<double>(wmax) / safe;
            npos[e] = per[e];
            nper[e] = per[e] / rate;
            nneg[e] = nper[e];
            UMAPHIP_REQUIRE(heads[e] >= 0 && heads[e] < h && tails[e] >= 0 && tails[e] < n,
                            UMAPHIP_NUMERICAL, "layout edge endpoint out of range");
            if (w > 0.0f) ++counts[heads[e]];
        }
        // build_head_csr: positive edges grouped by head, in edge-id order
        std::vector<int32_t> offsets(static_cast<size_t>(h + 1), 0);
        for (int64_t i = 0; i < h; ++i) offsets[i + 1] = offsets[i] + counts[i];
        std::vector<int32_t> ids(static_cast<size_t>(std::max<int64_t>(offsets[h], 1)), 1);
        std::vector<int32_t> cursor(offsets.begin(), offsets.end() - 1);
        for (int64_t e = 0; e < E; ++e)
            if (pruned[e] > 0.0f) ids[cursor[heads[e]]++] = static_cast<int32_t>(e);
        // optimizer_rng_from_first_coordinates
        StdRng rng(seed);
        int64_t base[3];
        for (int k = 0; k < 3; ++k)
            base[k] = static_cast<int64_t>(static_cast<int32_t>(0x80000000u + rng.below(0xFFFFFFFFu)));
        std::vector<int64_t> state(static_cast<size_t>(3 * h));
        for (int64_t i = 0; i < h; ++i) {
            uint64_t bits;
            const double first = static_cast<double>(head[i * d]);
            std::memcpy(&bits, &first, sizeof(bits));
            for (int k = 0; k < 3; ++k)
                state[3 * i + k] = static_cast<int64_t>(static_cast<uint64_t>(base[k]) + bits);
        }

        DeviceArray<float> dpos(static_cast<size_t>(h * d)), dtail, dw(std::max<int64_t>(E, 1));
        DeviceArray<int32_t> dtails(std::max<int64_t>(E, 1)), dmask(std::max<int64_t>(E, 1)),
            doff(h + 1), dids(ids.size());
        DeviceArray<double> dper(std::max<int64_t>(E, 1)), dnpos(std::max<int64_t>(E, 1)),
            dnper(std::max<int64_t>(E, 1)), dnneg(std::max<int64_t>(E, 1));
        DeviceArray<int64_t> drng(static_cast<size_t>(3 * h));
        hipStream_t s;
        HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard{s};
        auto up = [&](void *dst, const void *src, size_t bytes) {
            if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        };
        up(dpos.ptr, head, h * d * sizeof(float));
        if (!shared) {
            dtail.alloc(static_cast<size_t>(n * d));
            up(dtail.ptr, tail, n * d * sizeof(float));
        }
        up(dw.ptr, pruned.data(), E * sizeof(float));
        up(dtails.ptr, tails, E * sizeof(int32_t));
        up(dper.ptr, per.data(), E * sizeof(double));
        up(dnpos.ptr, npos.data(), E * sizeof(double));
        up(dnper.ptr, nper.data(), E * sizeof(double));
        up(dnneg.ptr, nneg.data(), E * sizeof(double));
        up(doff.ptr, offsets.data(), offsets.size() * sizeof(int32_t));
        up(dids.ptr, ids.data(), ids.size() * sizeof(int32_t));
        up(drng.ptr, state.data(), state.size() * sizeof(int64_t));
        ClassicArgs A{};
        A.h = h; A.n = n; A.d = d; A.E = E;
        A.pos = dpos.ptr; A.tail = shared ? dpos.ptr : dtail.ptr;
        A.tails = dtails.ptr; A.weights = dw.ptr;
        A.per = dper.ptr; A.npos = dnpos.ptr; A.nper = dnper.ptr; A.nneg = dnneg.ptr;
        A.mask = dmask.ptr; A.rng = drng.ptr;
        A.head_offsets = doff.ptr; A.head_edges = dids.ptr;
        A.a = a; A.b = b; A.gamma = gamma; A.rate = static_cast<float>(negative_sample_rate);
        A.shared = shared != 0;
        float alpha = learning_rate;
        for (int32_t t = 0; t < n_epochs; ++t) {
            A.epoch = static_cast<uint32_t>(t);
            A.alpha = alpha;
            classic_epoch_launch(A, s);
            alpha = learning_rate *
                    (1.0f - static_cast<float>(t) / static_cast<float>(std::max<int32_t>(n_epochs, 1)));
        }
        HIP_CHECK(hipMemcpyAsync(head, dpos.ptr, h * d * sizeof(float), hipMemcpyDeviceToHost, s));
        if (next_positive_out && E)
            HIP_CHECK(hipMemcpyAsync(next_positive_out, dnpos.ptr, E * sizeof(double),
                                     hipMemcpyDeviceToHost, s));
        if (next_negative_out && E)
            HIP_CHECK(hipMemcpyAsync(next_negative_out, dnneg.ptr, E * sizeof(double),
                                     hipMemcpyDeviceToHost, s));
        if (rng_out)
            HIP_CHECK(hipMemcpyAsync(rng_out, drng.ptr, 3 * h * sizeof(int64_t),
                                     hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    });
}

}  // extern "C"
