// C-ABI entry points of the modern layout and its device-resident driver.
//
// The driver restates modern_runtime::optimize_compact_inner
// (crates/umap/src/standard_umap/modern_runtime.rs:1037-1378) for one HIP
// stream: every input of the epoch loop that does not depend on the
// embedding (schedules, bias corrections, projection directions, adaptation
// sample indices, shuffled orders) is produced on the host up front or by a
// pipelined host step, so the 200-epoch loop is enqueued without any
// device-to-host synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "host_parallel.hpp"
#include "layout_host.hpp"
#include "layout_kernels.hpp"
#include "stdrng.hpp"

namespace umaphip {

namespace {
thread_local std::string g_last_error;
}

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

// ---------------------------------------------------------- host helpers
// adam_schedule / momentum_schedule (umap-core/src/modern_layout.rs:110-165),
// converted once to f32 (Schedule64::to_f32, :98-107).
void schedules(int epochs, double alpha0, bool good, bool momentum, std::vector<float> &alpha,
               std::vector<float> &mu) {
    alpha.resize(epochs);
    mu.resize(epochs);
    if (!momentum) {
        const int warm = good ? std::min(epochs / 8, 100) : std::min(epochs / 2, 100);
        for (int n = 0; n < epochs; ++n) {
            double a;
            if (n < warm) {
                const double k = good ? 0.5 : 2.0;
                const double r = 1.0 - static_cast<double>(n) / warm;
                a = (k * alpha0 - 0.1) * (r * r) + 0.1;
            } else {
                a = 0.15 * (1.0 - static_cast<double>(n - warm) / (epochs - warm)) + 0.05;
            }
            alpha[n] = static_cast<float>(a);
            mu[n] = 0.0f;
        }
    } else {
        const int warm = std::min(epochs, std::max(200, epochs / 8));
        for (int n = 0; n < epochs; ++n) {
            double r;
            if (good) {
                const double t = 1.0 - static_cast<double>(n) / warm;
                r = n < warm ? t * t : 0.0;
            } else {
                const double t = 1.0 - static_cast<double>(n) / epochs;
                r = 0.25 * (t * t);
            }
            alpha[n] = static_cast<float>(good ? 0.25 * alpha0 * r + 0.005 : alpha0 * r);
            mu[n] = static_cast<float>(0.5 * (1.0 - r));
        }
    }
}

uint64_t effective_range(uint64_t raw, uint64_t refs) {
    return std::min<uint64_t>(refs, std::max<uint64_t>(raw, 200));
}

// resolved_negative_scale (umap-core/src/modern_layout.rs:537-561).
float resolved_scale(uint64_t raw, uint64_t refs, bool has_explicit, float explicit_scale,
                     bool adaptation) {
    if (has_explicit) return explicit_scale;
    if (adaptation)
        return static_cast<float>(
            std::pow(static_cast<double>(std::min(raw, refs)) / static_cast<double>(refs), 0.75));
    return static_cast<float>(effective_range(raw, refs)) / static_cast<float>(refs);
}

// ProjectionOrder direction (modern_negatives.rs:208-218, 283-297).
void projection_direction(StdRng &rng, int64_t d, float *dir) {
    double sq = 0.0;
    for (int64_t c = 0; c < d; c += 2) {
        double g1, g2;
        rng.gaussian_pair(g1, g2);
        dir[c] = static_cast<float>(g1);
        sq += static_cast<double>(dir[c]) * static_cast<double>(dir[c]);
        if (c + 1 < d) {
            dir[c + 1] = static_cast<float>(g2);
            sq += static_cast<double>(dir[c + 1]) * static_cast<double>(dir[c + 1]);
        }
    }
    const double len = std::sqrt(sq);
    if (!std::isfinite(len) || len == 0.0) {
        std::fill(dir, dir + d, 0.0f);
        dir[0] = 1.0f;
    } else {
        for (int64_t c = 0; c < d; ++c) dir[c] = static_cast<float>(dir[c] / len);
    }
}

void bias_corrections(int epochs, std::vector<float> &c1, std::vector<float> &c2) {
    c1.resize(epochs);
    c2.resize(epochs);
    for (int t = 0; t < epochs; ++t) {
        c1[t] = static_cast<float>(1.0 - std::pow(static_cast<double>(0.9f), t + 1.0));
        c2[t] = static_cast<float>(1.0 - std::pow(static_cast<double>(0.99f), t + 1.0));
    }
}

namespace {

// Device bytes a plan gives its clock schedule (one byte per edge and epoch):
// 26.3M edges hold 163 epochs per pass.
constexpr int64_t kScheduleBytes = int64_t(4) << 30;

void check_layout_params(const umaphip_layout_params *p, int64_t h, int64_t n, int64_t d,
                         int64_t E) {
    UMAPHIP_REQUIRE(p != nullptr, UMAPHIP_INVALID_INPUT, "layout params are null");
    UMAPHIP_REQUIRE(h > 0 && n > 0 && d > 0, UMAPHIP_INVALID_INPUT,
                    "layout requires positive rows, references and components");
    UMAPHIP_REQUIRE(h <= INT32_MAX && n <= INT32_MAX && E <= INT32_MAX, UMAPHIP_OVERFLOW,
                    "layout sizes must fit i32");
    UMAPHIP_REQUIRE(p->optimizer == 0 || p->optimizer == 1, UMAPHIP_INVALID_CONFIG,
                    "optimizer must be adam or momentum");
    UMAPHIP_REQUIRE(p->epochs >= 0, UMAPHIP_INVALID_CONFIG, "epochs must be non-negative");
    UMAPHIP_REQUIRE(p->negative_selection_range > 0 && p->negative_selection_range <= INT32_MAX,
                    UMAPHIP_INVALID_CONFIG, "negative_selection_range must be a positive i32");
    UMAPHIP_REQUIRE(!p->has_scale || (std::isfinite(p->scale) && p->scale >= 0.0f),
                    UMAPHIP_INVALID_CONFIG, "negative sample scale must be finite and nonnegative");
    UMAPHIP_REQUIRE(p->adaptation_samples >= 0 && p->adaptation_samples <= INT32_MAX,
                    UMAPHIP_INVALID_CONFIG, "adaptation samples must fit i32");
    UMAPHIP_REQUIRE(p->negative_sample_rate >= 0, UMAPHIP_INVALID_CONFIG,
                    "negative sample rate must be nonnegative");
}

// Work partition of the unit sweep: unit u covers CSR edges
// [u*UE, (u+1)*UE). Rows are indexed among the rows with at least one edge.
struct UnitTables {
    int64_t units = 0;
    std::vector<int32_t> unit_k;     // [units][2] first/last nonzero-degree row index
    std::vector<int32_t> unit_frag;  // [units][2] fragment slot of the first/last row, or -1
    std::vector<int32_t> nz_rows;    // empty when every row has an edge
    std::vector<int32_t> nz_off;     // CSR start of each nonzero-degree row, then E
    std::vector<int32_t> row_split;  // [h] split-row index, or -1
    std::vector<int32_t> split_off;  // [splits+1] first fragment of each split row
};

UnitTables build_unit_tables(const uint32_t *offsets, int64_t h, int64_t E, int64_t UE) {
    UnitTables T;
    std::vector<int32_t> rows;
    rows.reserve(static_cast<size_t>(h));
    for (int64_t i = 0; i < h; ++i)
        if (offsets[i + 1] > offsets[i]) rows.push_back(static_cast<int32_t>(i));
    const int64_t K = static_cast<int64_t>(rows.size());
    T.nz_off.resize(static_cast<size_t>(K + 1));
    for (int64_t k = 0; k < K; ++k) T.nz_off[k] = static_cast<int32_t>(offsets[rows[k]]);
    T.nz_off[K] = static_cast<int32_t>(E);
    if (K != h) T.nz_rows = rows;
    T.units = (E + UE - 1) / UE;
    T.unit_k.resize(static_cast<size_t>(2 * T.units));
    T.unit_frag.assign(static_cast<size_t>(2 * T.units), -1);
    int64_t k = 0;
    for (int64_t u = 0; u < T.units; ++u) {
        const int64_t e0 = u * UE, last = std::min(E, e0 + UE) - 1;
        while (T.nz_off[k + 1] <= e0) ++k;
        int64_t k1 = k;
        while (T.nz_off[k1 + 1] <= last) ++k1;
        T.unit_k[2 * u] = static_cast<int32_t>(k);
        T.unit_k[2 * u + 1] = static_cast<int32_t>(k1);
    }
    T.row_split.assign(static_cast<size_t>(h), -1);
    T.split_off.push_back(0);
    int64_t frags = 0;
    for (int64_t q = 0; q < K; ++q) {
        const int64_t ua = T.nz_off[q] / UE, ub = (T.nz_off[q + 1] - 1) / UE;
        if (ua == ub) continue;
        T.row_split[rows[q]] = static_cast<int32_t>(T.split_off.size() - 1);
        for (int64_t w = ua; w <= ub; ++w) {
            const int32_t slot = static_cast<int32_t>(frags + (w - ua));
            if (T.unit_k[2 * w] == q) T.unit_frag[2 * w] = slot;
            else T.unit_frag[2 * w + 1] = slot;
        }
        frags += ub - ua + 1;
        UMAPHIP_REQUIRE(frags <= INT32_MAX, UMAPHIP_OVERFLOW, "too many row fragments");
        T.split_off.push_back(static_cast<int32_t>(frags));
    }
    return T;
}

}  // namespace
}  // namespace umaphip

using namespace umaphip;

// =================================================================== plan
struct umaphip_plan {
    int64_t h = 0, n = 0, d = 0, E = 0;
    bool shared = true, momentum = false, hard = false, small = false, alias = false;
    bool sampling = false;
    umaphip_layout_params p{};
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;

    // state
    DeviceArray<float> pos[2], head0, tail, updates, m, v;
    DeviceArray<int32_t> offsets, indices;
    DeviceArray<double> periods, negperiods, npos, nneg, npos0, nneg0;
    // negatives
    DeviceArray<float> table, sorted_vals, query_proj, dirs;
    DeviceArray<uint32_t> keys, keys_sorted;
    DeviceArray<int32_t> iota, proj_order, ranks, host_orders[2], source_dev[2], block_of_row;
    DeviceArray<uint8_t> sort_ws;
    DeviceArray<unsigned long long> status;
    // adaptation
    DeviceArray<int32_t> samples_dev;      // per adaptation: S + 5S + 5S
    DeviceArray<double> resultants, scale_state;
    DeviceArray<float> scale_dev;
    DeviceArray<int32_t> error_flag;
    std::vector<int32_t> adaptation_epochs;
    int64_t sample_count = 0;

    // optional per-kernel timing (events around every edge sweep / refresh)
    bool profiling = false;
    std::vector<hipEvent_t> sweep_ev, refresh_ev;
    int sweep_used = 0, refresh_used = 0;

    // optional FP64 objective trace (every obj_every epochs and at the end)
    int32_t obj_every = 0;
    int64_t obj_count = 0;
    DeviceArray<float> obj_weights;
    DeviceArray<double> obj_ws, obj_vals;
    std::vector<int32_t> obj_epochs;

    // edge-balanced unit sweep (Adam, d = 2, H > 1024)
    bool unit_path = false;
    bool compact = false;   // 24-byte records (rank-ordered plans whose periods allow it)
    bool compactable = false;  // every period is rebuilt exactly from its f32 sample count
    int64_t units = 0;
    DeviceArray<double> rec;
    DeviceArray<uint16_t> fire;
    DeviceArray<int32_t> unit_k, unit_frag, nz_rows, nz_off, row_split, split_off;
    DeviceArray<float> grad, frag;
    DeviceArray<uint8_t> split;   // rank-ordered sweep: row split over units (set by the sweep)
    // rank-ordered unit sweep (projection windows, every row with an edge)
    bool ranked = false;
    int64_t ranked_epochs = 0;
    int32_t row_cap_id = 0, row_cap_ranked = 0;   // unit row-table capacity per order
    DeviceArray<int32_t> pref, pref_blocks, rank_unit_k, rank_seg, rank_deg;
    // {x, y, CSR start, degree} per row, kept by the Adam step for the
    // refresh's last sort pass (rank-ordered plans)
    DeviceArray<float> rows4;
    // Clock schedule of the edge-balanced sweep: each edge's term count per
    // epoch, computed sched_span epochs at a time (0: off, the sweep
    // advances the clocks from the records every epoch).
    bool schedulable = false;
    int32_t sched_span = 0;
    int64_t sched_stride = 0;
    DeviceArray<uint8_t> sched;

    // optional trace of one unit-sweep epoch: the state it starts from
    // (positions, negative order, ranks, scale), each row's assembled
    // gradient and every negative drawn per edge
    int32_t trace_epoch = -1, trace_cap = 0;
    bool traced = false;
    int32_t trace_info[3] = {0, 0, 0};   // rank-ordered, compact records, row-table capacity
    DeviceArray<float> tr_head, tr_grad, tr_scale;
    DeviceArray<int32_t> tr_order, tr_ranks, tr_count, tr_cand, tr_misc;  // misc: overflow, paths[2]
    DeviceArray<uint8_t> tr_kept;

    std::vector<float> alpha, mu, c1, c2;
    float scale0 = 1.0f;
    uint32_t range = 0;
    int64_t refresh_count = 0;
    int cur = 0;  // which pos buffer holds the current head after a run
    bool ran = false;

    ~umaphip_plan() {
        for (auto e : sweep_ev) (void)hipEventDestroy(e);
        for (auto e : refresh_ev) (void)hipEventDestroy(e);
        if (ev_start) (void)hipEventDestroy(ev_start);
        if (ev_stop) (void)hipEventDestroy(ev_stop);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int64_t adaptation_stride() const { return sample_count * 11; }

    // cap < 0: automatic span (every epoch in one pass when the slabs fit
    // kScheduleBytes); 0: off; > 0: at most cap epochs per pass.
    void configure_schedule(int32_t cap) {
        if (!schedulable || cap == 0) {
            sched_span = 0;
            sched.alloc(0);
            return;
        }
        sched_stride = unit_schedule_stride(E);
        int64_t span = std::max<int64_t>(std::min<int64_t>(p.epochs, kScheduleBytes / sched_stride), 1);
        if (cap > 0) span = std::min<int64_t>(span, cap);
        sched_span = static_cast<int32_t>(span);
        try {
            sched.alloc(static_cast<size_t>(sched_stride * span));
        } catch (const Fail &f) {
            // the automatic schedule is an optimisation: without the memory
            // the sweep advances the clocks itself
            sched_span = 0;
            if (cap > 0 || f.status != UMAPHIP_MEMORY) throw;
            (void)hipGetLastError();
            set_error("%s", "");
        }
    }

    void mark(std::vector<hipEvent_t> &pool, int &used) {
        if (!profiling) return;
        if (used >= static_cast<int>(pool.size())) {
            hipEvent_t e;
            HIP_CHECK(hipEventCreate(&e));
            pool.push_back(e);
        }
        HIP_CHECK(hipEventRecord(pool[used++], stream));
    }

    // Sum of the durations between consecutive event pairs.
    double pairs_ms(std::vector<hipEvent_t> &pool, int used) {
        double total = 0.0;
        for (int i = 0; i + 1 < used; i += 2) {
            float ms = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&ms, pool[i], pool[i + 1]));
            total += ms;
        }
        return total;
    }

    // One projection refresh (ProjectionOrder::refresh, modern_negatives.rs:
    // 283-351): project -> stable argsort -> ranks (+ negative table).
    void refresh(int64_t k, const float *head_now) {
        mark(refresh_ev, refresh_used);
        refresh_body(k, head_now);
        mark(refresh_ev, refresh_used);
    }

    void refresh_body(int64_t k, const float *head_now) {
        const float *dir = dirs.ptr + k * d;
        const float *reference = shared ? head_now : tail.ptr;
        const unsigned long long tag = static_cast<unsigned long long>(k) << 32;
        if (shared && d == 2) {
            // projection, its first histogram and the four sort passes; the
            // last pass writes the ranks and the table
            // (rank-ordered sweeps: also each rank's CSR start and degree)
            // (rank-ordered plans: from the packed rows the Adam step keeps,
            // and no ranks, which only the epoch trace reads)
            launch_project_sort_ranked(sort_ws.ptr, sort_ws.count, head_now, n, dir, status.ptr, tag,
                                       keys.ptr, keys_sorted.ptr, proj_order.ptr,
                                       ranked ? nullptr : ranks.ptr,
                                       table.ptr, stream, ranked ? offsets.ptr : nullptr,
                                       ranked ? rank_seg.ptr : nullptr,
                                       ranked ? rank_deg.ptr : nullptr, error_flag.ptr,
                                       ranked ? rows4.ptr : nullptr);
            return;
        }
        launch_project_keys(reference, n, d, dir, keys.ptr, nullptr, status.ptr, tag, stream);
        if (shared) {
            // the sort's last pass writes the ranks and (d = 2) the table
            launch_sort_pairs_ranked(sort_ws.ptr, sort_ws.count, keys.ptr, keys_sorted.ptr, nullptr,
                                     proj_order.ptr, n, ranks.ptr, d == 2 ? head_now : nullptr,
                                     d == 2 ? table.ptr : nullptr, stream);
        } else {
            launch_sort_pairs(sort_ws.ptr, sort_ws.count, keys.ptr, keys_sorted.ptr, nullptr,
                              proj_order.ptr, n, stream);
            launch_values_from_keys(keys_sorted.ptr, n, sorted_vals.ptr, stream);
            launch_project_keys(head_now, h, d, dir, nullptr, query_proj.ptr, status.ptr, tag,
                                stream);
            launch_lower_bound(sorted_vals.ptr, n, query_proj.ptr, h, ranks.ptr, stream);
            if (d == 2) launch_gather_table(proj_order.ptr, n, tail.ptr, table.ptr, stream);
        }
    }

    void run();
};

void umaphip_plan::run() {
    const int epochs = p.epochs;
    const bool adam = !momentum;
    ranked_epochs = 0;
    // reset state
    HIP_CHECK(hipMemcpyAsync(pos[0].ptr, head0.ptr, head0.bytes(), hipMemcpyDeviceToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(npos.ptr, npos0.ptr, npos0.bytes(), hipMemcpyDeviceToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(nneg.ptr, nneg0.ptr, nneg0.bytes(), hipMemcpyDeviceToDevice, stream));
    if (unit_path)
        launch_unit_pack(npos0.ptr, nneg0.ptr, periods.ptr, indices.ptr, E, range, compact,
                         static_cast<double>(p.epochs), rec.ptr, fire.ptr, stream);
    if (ranked) launch_pack_rows(pos[0].ptr, offsets.ptr, h, rows4.ptr, stream);
    HIP_CHECK(hipMemsetAsync(updates.ptr, 0, updates.bytes(), stream));
    if (adam) {
        HIP_CHECK(hipMemsetAsync(m.ptr, 0, m.bytes(), stream));
        HIP_CHECK(hipMemsetAsync(v.ptr, 0, v.bytes(), stream));
    }
    const double state0[4] = {static_cast<double>(scale0), static_cast<double>(scale0), 0.0, 0.0};
    HIP_CHECK(hipMemcpyAsync(scale_state.ptr, state0, sizeof(state0), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(scale_dev.ptr, &scale0, sizeof(float), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemsetAsync(error_flag.ptr, 0, sizeof(int32_t), stream));
    launch_status_min_u64(status.ptr, ~0ull, stream);
    HIP_CHECK(hipStreamSynchronize(stream));  // host staging above is pageable

    sweep_used = 0;
    refresh_used = 0;
    obj_count = 0;
    traced = false;
    if (trace_epoch >= 0) {
        HIP_CHECK(hipMemsetAsync(tr_count.ptr, 0xFF, tr_count.bytes(), stream));   // -1: no fire
        HIP_CHECK(hipMemsetAsync(tr_misc.ptr, 0, tr_misc.bytes(), stream));
    }
    HIP_CHECK(hipEventRecord(ev_start, stream));
    OrderReplay orders(h, n, p.seed, alias, hard, momentum);
    const bool upload_negative = !hard;  // global mode reads host orders
    const bool upload_source = momentum;
    std::unique_ptr<Pinned> stage_neg[2], stage_src[2];
    hipEvent_t staged[2] = {nullptr, nullptr};
    if (upload_negative || upload_source) {
        for (int k = 0; k < 2; ++k) {
            if (upload_negative) stage_neg[k].reset(new Pinned(sizeof(int32_t) * n));
            if (upload_source) stage_src[k].reset(new Pinned(sizeof(int32_t) * h));
            HIP_CHECK(hipEventCreateWithFlags(&staged[k], hipEventDisableTiming));
        }
    }
    struct EventGuard {
        hipEvent_t *e;
        ~EventGuard() {
            for (int k = 0; k < 2; ++k)
                if (e[k]) (void)hipEventDestroy(e[k]);
        }
    } guard{staged};

    int c = 0;  // pos[c] is the current head
    int64_t refreshes = 0;
    bool projection_ready = false;
    if (hard && h > 1024) {
        refresh(refreshes++, pos[c].ptr);
        projection_ready = true;
    }
    size_t next_adapt = 0;
    const int64_t block = h <= 1024 ? h : std::max<int64_t>(h / 2, 1);

    for (int t = 0; t < epochs; ++t) {
        const int slot = t & 1;
        // ---- orders for this epoch (host replay, pinned upload)
        const int32_t *neg_ids = nullptr;
        const int32_t *src_ids = nullptr;
        if (upload_negative || upload_source) {
            if (t >= 2) HIP_CHECK(hipEventSynchronize(staged[slot]));
            if (upload_negative) {
                std::memcpy(stage_neg[slot]->ptr, orders.neg().data(), sizeof(int32_t) * n);
                HIP_CHECK(hipMemcpyAsync(host_orders[slot].ptr, stage_neg[slot]->ptr,
                                         sizeof(int32_t) * n, hipMemcpyHostToDevice, stream));
                neg_ids = host_orders[slot].ptr;
            }
            if (upload_source) {
                std::memcpy(stage_src[slot]->ptr, orders.source.data(), sizeof(int32_t) * h);
                HIP_CHECK(hipMemcpyAsync(source_dev[slot].ptr, stage_src[slot]->ptr,
                                         sizeof(int32_t) * h, hipMemcpyHostToDevice, stream));
                src_ids = source_dev[slot].ptr;
            }
            HIP_CHECK(hipEventRecord(staged[slot], stream));
        }
        // negatives in force this epoch
        const float *src_pos = shared ? pos[c].ptr : tail.ptr;
        const int32_t *order_now;
        const float *table_now = nullptr;
        if (projection_ready) {
            order_now = proj_order.ptr;
            if (d == 2) table_now = table.ptr;
        } else if (hard) {
            order_now = iota.ptr;  // identity order before the first refresh
            if (d == 2) table_now = src_pos;
        } else {
            order_now = neg_ids;
            if (d == 2 && adam) {
                launch_gather_table(order_now, n, src_pos, table.ptr, stream);
                table_now = table.ptr;
            }
        }
        // ---- adaptation before the gradient (modern_runtime.rs:1192-1226)
        if (sampling && next_adapt < adaptation_epochs.size() && adaptation_epochs[next_adapt] == t) {
            const int32_t *base = samples_dev.ptr + next_adapt * adaptation_stride();
            launch_resultants(pos[c].ptr, n, d, order_now, offsets.ptr, indices.ptr, base,
                              base + sample_count, base + 6 * sample_count, sample_count,
                              static_cast<uint32_t>(p.negative_selection_range), p.a, p.b,
                              p.repulsion, p.exclude_graph_neighbors, resultants.ptr, stream);
            launch_ratio_update(resultants.ptr, sample_count, scale_state.ptr, scale_dev.ptr,
                                error_flag.ptr, stream);
            ++next_adapt;
        }
        // ---- the epoch (modern_epoch, kernels/.../modern_layout.rs:925-1026)
        if (adam && d == 2) {
            SweepArgs A{};
            A.h = h; A.n = n;
            A.pos = pos[c].ptr; A.pos_out = pos[1 - c].ptr;
            A.tgt = src_pos; A.table = table_now; A.order = order_now;
            A.ranks = ranks.ptr;
            A.offsets = offsets.ptr; A.indices = indices.ptr;
            A.periods = periods.ptr; A.next_positive = npos.ptr;
            A.negative_periods = negperiods.ptr; A.next_negative = nneg.ptr;
            A.m = m.ptr; A.v = v.ptr;
            A.scale_ptr = scale_dev.ptr;
            A.epoch = static_cast<uint32_t>(t);
            A.a = p.a; A.b = p.b; A.gamma = p.repulsion; A.range = range;
            A.multiplier = 1.0f; A.alpha = alpha[t]; A.c1 = c1[t]; A.c2 = c2[t];
            A.shared = shared; A.negatives = p.negative_sample_rate != 0;
            A.exclude = p.exclude_graph_neighbors; A.fused = 1;
            mark(sweep_ev, sweep_used);   // schedule passes count as sweep time
            const uint8_t *sched_now = nullptr;
            if (unit_path && sched_span > 0) {
                const int slab = t % sched_span;
                if (slab == 0)
                    launch_unit_schedule(rec.ptr, E, compact, static_cast<double>(p.epochs),
                                         static_cast<double>(p.negative_sample_rate),
                                         p.negative_sample_rate != 0, t,
                                         std::min(sched_span, epochs - t), sched.ptr, sched_stride,
                                         error_flag.ptr, stream);
                sched_now = sched.ptr + slab * sched_stride;
            }
            if (unit_path) {
                UnitSweepArgs U{};
                U.s = A;
                U.rec = rec.ptr; U.fire = fire.ptr;
                U.compact = compact;
                U.per_num = static_cast<double>(p.epochs);
                U.sched = sched_now;
                U.unit_k = unit_k.ptr; U.unit_frag = unit_frag.ptr;
                U.nz_rows = nz_rows.count ? nz_rows.ptr : nullptr; U.nz_off = nz_off.ptr;
                U.grad = grad.ptr; U.frag = frag.ptr; U.split = split.ptr;
                U.units = units; U.n_edges = E;
                U.rate = static_cast<double>(p.negative_sample_rate);
                // with this epoch's projection order, sweep the rows by rank
                const bool by_rank = ranked && projection_ready;
                // Without the clock schedule, the rank-ordered sweep decides
                // firing from next_positive in the record and leaves the
                // firing shadow behind; the id-order sweep reads the shadow,
                // so it cannot follow a ranked epoch.
                UMAPHIP_REQUIRE(by_rank || ranked_epochs == 0, UMAPHIP_BACKEND,
                                "unit sweep: id-order epoch after a rank-ordered epoch");
                U.row_cap = by_rank ? row_cap_ranked : row_cap_id;
                if (by_rank) {
                    launch_rank_prefix(rank_deg.ptr, n, static_cast<int32_t>(E), pref.ptr,
                                       pref_blocks.ptr, rank_unit_k.ptr, stream);
                    U.pref = pref.ptr;
                    U.rank_seg = rank_seg.ptr;
                    U.unit_k = rank_unit_k.ptr;
                    ++ranked_epochs;
                }
                const bool tracing = t == trace_epoch;
                if (tracing) {
                    const hipMemcpyKind dd = hipMemcpyDeviceToDevice;
                    // rank-ordered refreshes leave the ranks unwritten
                    if (ranked) launch_inverse_ranks(order_now, n, ranks.ptr, stream);
                    HIP_CHECK(hipMemcpyAsync(tr_head.ptr, pos[c].ptr, tr_head.bytes(), dd, stream));
                    HIP_CHECK(hipMemcpyAsync(tr_order.ptr, order_now, tr_order.bytes(), dd, stream));
                    HIP_CHECK(hipMemcpyAsync(tr_ranks.ptr, hard ? ranks.ptr : iota.ptr,
                                             tr_ranks.bytes(), dd, stream));
                    HIP_CHECK(hipMemcpyAsync(tr_scale.ptr, scale_dev.ptr, sizeof(float), dd, stream));
                    U.tr_count = tr_count.ptr;
                    U.tr_cand = tr_cand.ptr;
                    U.tr_kept = tr_kept.ptr;
                    U.tr_cap = trace_cap;
                    U.tr_overflow = tr_misc.ptr;
                    U.tr_paths = tr_misc.ptr + 1;
                    trace_info[0] = by_rank ? 1 : 0;
                    trace_info[1] = compact ? 1 : 0;
                    trace_info[2] = U.row_cap;
                    traced = true;
                }
                float *grad_out = tracing ? tr_grad.ptr : nullptr;
                launch_unit_sweep(U, stream);
                if (by_rank)
                    launch_unit_adam_ranked(h, pos[c].ptr, pos[1 - c].ptr, grad.ptr, split.ptr,
                                            frag.ptr, m.ptr, v.ptr, alpha[t], c1[t], c2[t], stream,
                                            grad_out, offsets.ptr, rows4.ptr);
                else
                    launch_unit_adam(h, pos[c].ptr, pos[1 - c].ptr, grad.ptr, row_split.ptr,
                                     split_off.ptr, frag.ptr, m.ptr, v.ptr, alpha[t], c1[t],
                                     c2[t], stream, grad_out);
            } else {
                launch_edge_sweep(A, stream);
            }
            mark(sweep_ev, sweep_used);
            c = 1 - c;
        } else {
            const float scale_host = scale0;  // momentum never adapts the scale
            RowwiseArgs R{};
            R.scale_ptr = scale_dev.ptr;
            R.h = h; R.n = n; R.d = d;
            R.head = pos[c].ptr; R.hs0 = d; R.hs1 = 1;
            R.tail = shared ? pos[c].ptr : tail.ptr; R.ts0 = d; R.ts1 = 1;
            R.offsets = offsets.ptr; R.indices = indices.ptr;
            R.source_order = src_ids ? src_ids : iota.ptr;
            R.negative_order = order_now; R.source_ranks = projection_ready ? ranks.ptr : iota.ptr;
            R.periods = periods.ptr; R.next_positive = npos.ptr;
            R.negative_periods = negperiods.ptr; R.next_negative = nneg.ptr;
            R.updates = updates.ptr; R.us0 = d; R.us1 = 1;
            R.epoch = static_cast<uint32_t>(t);
            R.a = p.a; R.b = p.b; R.gamma = p.repulsion; R.scale = scale_host; R.range = range;
            R.shared = shared; R.negatives = p.negative_sample_rate != 0;
            R.exclude = p.exclude_graph_neighbors;
            if (adam) {
                R.multiplier = 1.0f;
                launch_gradient_rowwise(R, 0, h, stream);
                launch_adam(h, d, pos[c].ptr, d, 1, updates.ptr, d, 1, m.ptr, d, 1, v.ptr, d, 1,
                            alpha[t], c1[t], c2[t], stream);
            } else {
                R.multiplier = alpha[t];
                const int32_t *so = src_ids ? src_ids : iota.ptr;
                if (d == 2) launch_block_of_row(so, h, block, block_of_row.ptr, stream);
                int32_t blk = 0;
                for (int64_t start = 0; start < h; start += block, ++blk) {
                    const int64_t end = std::min(start + block, h);
                    if (d == 2) {
                        SweepArgs A{};
                        A.h = h; A.n = n;
                        A.pos = pos[c].ptr; A.tgt = shared ? pos[c].ptr : tail.ptr;
                        A.order = order_now; A.ranks = projection_ready ? ranks.ptr : iota.ptr;
                        A.offsets = offsets.ptr; A.indices = indices.ptr;
                        A.periods = periods.ptr; A.next_positive = npos.ptr;
                        A.negative_periods = negperiods.ptr; A.next_negative = nneg.ptr;
                        A.updates = updates.ptr;
                        A.block_of_row = block_of_row.ptr; A.block_id = blk;
                        A.scale = scale_host;
                        A.epoch = static_cast<uint32_t>(t);
                        A.a = p.a; A.b = p.b; A.gamma = p.repulsion; A.range = range;
                        A.multiplier = alpha[t];
                        A.shared = shared; A.negatives = p.negative_sample_rate != 0;
                        A.exclude = p.exclude_graph_neighbors; A.fused = 0;
                        launch_edge_sweep(A, stream);
                        launch_momentum_apply_rows(h, pos[c].ptr, updates.ptr, block_of_row.ptr,
                                                   blk, stream);
                    } else {
                        launch_gradient_rowwise(R, start, end, stream);
                        launch_momentum_apply(d, pos[c].ptr, d, 1, updates.ptr, d, 1, so, start,
                                              end, stream);
                    }
                }
                launch_momentum_decay(h, d, updates.ptr, d, 1, mu[t], stream);
            }
        }
        if (obj_every > 0 && ((t + 1) % obj_every == 0 || t + 1 == epochs)) {
            launch_cross_entropy(pos[c].ptr, h, shared ? pos[c].ptr : tail.ptr, d, offsets.ptr,
                                 indices.ptr, obj_weights.ptr, p.a, p.b, obj_ws.ptr,
                                 obj_vals.ptr + obj_count, stream);
            ++obj_count;
        }
        // ---- next orders, then the projection refresh
        orders.advance();
        if (hard) {
            if (t + 1 < epochs) {
                refresh(refreshes++, pos[c].ptr);
                projection_ready = true;
            } else {
                // The reference refreshes after the final epoch too; only its
                // finiteness check is observable.
                const unsigned long long tag = static_cast<unsigned long long>(refreshes) << 32;
                launch_project_keys(shared ? pos[c].ptr : tail.ptr, n, d,
                                    dirs.ptr + refreshes * d, nullptr, sorted_vals.ptr,
                                    status.ptr, tag, stream);
                if (!shared)
                    launch_project_keys(pos[c].ptr, h, d, dirs.ptr + refreshes * d, nullptr,
                                        query_proj.ptr, status.ptr, tag, stream);
                ++refreshes;
            }
        }
    }
    HIP_CHECK(hipEventRecord(ev_stop, stream));
    if (unit_path) launch_unit_unpack(rec.ptr, E, compact, npos.ptr, nneg.ptr, stream);
    cur = c;
    refresh_count = refreshes;
    ran = true;
}

extern "C" {

const char *umaphip_last_error(void) { return g_last_error.c_str(); }

int32_t umaphip_version(void) { return 100; }

int32_t umaphip_device_count(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return 0;
    return count;
}

umaphip_status umaphip_plan_create(int64_t h, int64_t n, int64_t d, const float *head,
                                   const float *tail, const uint32_t *offsets,
                                   const uint32_t *indices, int64_t E, const double *periods,
                                   const double *next_positive, const double *negative_periods,
                                   const double *next_negative,
                                   const umaphip_layout_params *params, int32_t shared,
                                   umaphip_plan **out) {
    return guarded([&] {
        UMAPHIP_REQUIRE(out != nullptr, UMAPHIP_INVALID_INPUT, "plan output pointer is null");
        *out = nullptr;
        if (shared) n = h;
        check_layout_params(params, h, n, d, E);
        UMAPHIP_REQUIRE(head && offsets && (E == 0 || (indices && periods && next_positive &&
                                                       negative_periods && next_negative)),
                        UMAPHIP_INVALID_INPUT, "layout arrays are null");
        UMAPHIP_REQUIRE(shared || tail, UMAPHIP_INVALID_INPUT, "transform requires a tail");
        UMAPHIP_REQUIRE(offsets[0] == 0 && static_cast<int64_t>(offsets[h]) == E,
                        UMAPHIP_INVALID_INPUT, "CSR offsets must start at 0 and end at E");
        for (int64_t i = 0; i < h; ++i)
            UMAPHIP_REQUIRE(offsets[i] <= offsets[i + 1], UMAPHIP_INVALID_INPUT,
                            "CSR offsets must be nondecreasing (row %lld)", (long long)i);
        for (int64_t e = 0; e < E; ++e)
            UMAPHIP_REQUIRE(static_cast<int64_t>(indices[e]) < n, UMAPHIP_INVALID_INPUT,
                            "CSR index %lld out of range", (long long)e);
        std::unique_ptr<umaphip_plan> P(new umaphip_plan());
        P->h = h; P->n = n; P->d = d; P->E = E;
        P->p = *params;
        P->shared = shared != 0;
        P->momentum = params->optimizer == 1;
        const int64_t Q = params->negative_selection_range;
        P->hard = Q < n;
        P->small = h <= 1024;
        P->alias = P->shared && P->momentum && !P->hard;
        P->sampling = P->hard && P->shared && !P->momentum && params->adaptation_samples > 0;
        P->range = static_cast<uint32_t>(effective_range(static_cast<uint64_t>(Q), n));
        P->scale0 = resolved_scale(static_cast<uint64_t>(Q), n, params->has_scale != 0,
                                   params->scale, params->adaptation_samples > 0);
        schedules(params->epochs, params->initial_alpha, params->good_initialization != 0,
                  P->momentum, P->alpha, P->mu);
        bias_corrections(params->epochs, P->c1, P->c2);
        HIP_CHECK(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreate(&P->ev_start));
        HIP_CHECK(hipEventCreate(&P->ev_stop));

        const size_t hd = static_cast<size_t>(h * d);
        P->head0.alloc(hd);
        P->pos[0].alloc(hd);
        P->pos[1].alloc(hd);
        P->updates.alloc(hd);
        if (!P->momentum) {
            P->m.alloc(hd);
            P->v.alloc(hd);
        }
        HIP_CHECK(hipMemcpy(P->head0.ptr, head, hd * sizeof(float), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(P->pos[1].ptr, head, hd * sizeof(float), hipMemcpyHostToDevice));
        if (!P->shared) {
            P->tail.alloc(static_cast<size_t>(n * d));
            HIP_CHECK(hipMemcpy(P->tail.ptr, tail, n * d * sizeof(float), hipMemcpyHostToDevice));
        }
        P->offsets.alloc(h + 1);
        HIP_CHECK(hipMemcpy(P->offsets.ptr, offsets, (h + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
        const size_t Ec = static_cast<size_t>(std::max<int64_t>(E, 1));
        P->indices.alloc(Ec);
        P->periods.alloc(Ec);
        P->negperiods.alloc(Ec);
        P->npos.alloc(Ec);
        P->nneg.alloc(Ec);
        P->npos0.alloc(Ec);
        P->nneg0.alloc(Ec);
        if (E) {
            HIP_CHECK(hipMemcpy(P->indices.ptr, indices, E * sizeof(int32_t), hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(P->periods.ptr, periods, E * sizeof(double), hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(P->negperiods.ptr, negative_periods, E * sizeof(double),
                                hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(P->npos0.ptr, next_positive, E * sizeof(double), hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(P->nneg0.ptr, next_negative, E * sizeof(double), hipMemcpyHostToDevice));
        }
        P->iota.alloc(std::max(h, n));
        launch_iota(P->iota.ptr, std::max(h, n), nullptr);
        P->ranks.alloc(std::max(h, n));
        P->table.alloc(static_cast<size_t>(n * 2));
        P->status.alloc(1);
        P->scale_state.alloc(4);
        P->scale_dev.alloc(1);
        P->error_flag.alloc(1);
        P->sorted_vals.alloc(n);
        P->query_proj.alloc(h);
        if (P->hard) {
            const int64_t nref = params->epochs + (h > 1024 ? 1 : 0);
            std::vector<float> dirs(static_cast<size_t>(std::max<int64_t>(nref, 1) * d));
            StdRng rng(params->seed);
            for (int64_t k = 0; k < nref; ++k) projection_direction(rng, d, dirs.data() + k * d);
            P->dirs.alloc(dirs.size());
            HIP_CHECK(hipMemcpy(P->dirs.ptr, dirs.data(), dirs.size() * sizeof(float),
                                hipMemcpyHostToDevice));
            P->keys.alloc(n);
            P->keys_sorted.alloc(n);
            P->proj_order.alloc(n);
            if (P->shared && d == 2) {
                P->sort_ws.alloc(project_sort_workspace(n));
                clear_project_sort_workspace(P->sort_ws.ptr, P->sort_ws.count, n, P->stream);
            } else {
                P->sort_ws.alloc(sort_pairs_workspace(n));
            }
        }
        if (!P->hard) {
            P->host_orders[0].alloc(n);
            P->host_orders[1].alloc(n);
        }
        // Edge-balanced unit sweep: Adam fits, d = 2, large layouts, epochs that
        // fit the 16-bit firing shadow, and negative periods equal to the
        // reference formula (period / rate) * 1.5 for every edge. Transform
        // rows have at most n_neighbors edges (no hub rows), so transforms
        // keep the tile sweep, whose per-row sums the transform server
        // reproduces bit for bit.
        bool unit_ok = P->shared && !P->momentum && d == 2 && !P->small && E > 0 &&
                       params->epochs < 65535;
        if (unit_ok && params->negative_sample_rate > 0) {
            const double rate = static_cast<double>(params->negative_sample_rate);
            std::atomic<int> differs{0};
            parallel_for(E, [&](int64_t e) {
                if (negative_periods[e] != (periods[e] / rate) * 1.5) differs.store(1);
            }, 1 << 16);
            unit_ok = !differs.load();
        }
        if (unit_ok) {
            UnitTables T = build_unit_tables(offsets, h, E, unit_edges());
            P->unit_path = true;
            P->units = T.units;
            P->fire.alloc(static_cast<size_t>(E));
            auto up = [](DeviceArray<int32_t> &dst, const std::vector<int32_t> &src) {
                if (src.empty()) return;
                dst.alloc(src.size());
                HIP_CHECK(hipMemcpy(dst.ptr, src.data(), src.size() * sizeof(int32_t),
                                    hipMemcpyHostToDevice));
            };
            up(P->unit_k, T.unit_k);
            up(P->unit_frag, T.unit_frag);
            up(P->nz_rows, T.nz_rows);
            up(P->nz_off, T.nz_off);
            up(P->row_split, T.row_split);
            up(P->split_off, T.split_off);
            P->grad.alloc(static_cast<size_t>(2 * h));
            HIP_CHECK(hipMemset(P->grad.ptr, 0, P->grad.bytes()));
            // Projection windows over a graph whose every row has an edge:
            // the epochs with a projection order sweep rows by rank (fragment
            // slots 2u and 2u + 1 of every unit).
            P->ranked = P->hard && T.nz_rows.empty() && E < INT32_MAX;
            if (P->ranked) {
                P->split.alloc(static_cast<size_t>(h));
                HIP_CHECK(hipMemset(P->split.ptr, 0, P->split.bytes()));
            }
            // Rank-ordered plans keep 24-byte records when every period is
            // rebuilt exactly from an f32 sample count.
            P->compactable = P->ranked &&
                             periods_compactable(periods, E, static_cast<double>(params->epochs));
            P->compact = P->compactable;
            P->rec.alloc(static_cast<size_t>((P->compact ? 3 : 4) * E));
            // The clocks run ahead of the sweep as a per-epoch schedule of
            // term counts (one byte per edge and epoch); a firing edge draws
            // at most about 1.34 rate + 3 negatives, so rate <= 64 keeps
            // every count in a byte.
            P->schedulable = params->negative_sample_rate <= 64;
            P->configure_schedule(-1);
            // Row-table capacity: the rows of a unit in CSR order are known;
            // in any rank order a unit of UE slots holds at most
            // 2 + (UE - 2) / (smallest degree) rows (two partial rows).
            const int64_t UE = unit_edges(), narrow = unit_narrow_rows();
            int64_t id_rows = 0, min_deg = E;
            for (int64_t u = 0; u < T.units; ++u)
                id_rows = std::max<int64_t>(id_rows, T.unit_k[2 * u + 1] - T.unit_k[2 * u] + 1);
            for (int64_t i = 0; i < h; ++i)
                min_deg = std::min<int64_t>(min_deg, offsets[i + 1] - offsets[i]);
            P->row_cap_id = static_cast<int32_t>(id_rows <= narrow ? narrow : UE);
            P->row_cap_ranked = static_cast<int32_t>(
                min_deg > 0 && 2 + (UE - 2) / min_deg <= narrow ? narrow : UE);
            int64_t frag_slots = std::max<int64_t>(T.split_off.back(), 1);
            if (P->ranked) {
                frag_slots = std::max<int64_t>(frag_slots, 2 * T.units);
                P->pref.alloc(static_cast<size_t>(n + 1));
                P->pref_blocks.alloc(static_cast<size_t>(rank_prefix_blocks(n)));
                P->rank_unit_k.alloc(static_cast<size_t>(2 * T.units));
                P->rank_seg.alloc(static_cast<size_t>(n));
                P->rank_deg.alloc(static_cast<size_t>(n));
                P->rows4.alloc(static_cast<size_t>(4 * h));
            }
            P->frag.alloc(static_cast<size_t>(2 * frag_slots));
        }
        if (P->momentum) {
            P->source_dev[0].alloc(h);
            P->source_dev[1].alloc(h);
            P->block_of_row.alloc(h);
        }
        if (P->sampling) {
            // ForceSamples (modern_negatives.rs:378-470): StdRng(42) draws S
            // sources, 5S local offsets, 5S global ranks per adaptation, in
            // adaptation order; the draws do not depend on the embedding.
            const int64_t S = std::min<int64_t>(params->adaptation_samples, n);
            UMAPHIP_REQUIRE(S <= kDeviceRatioMax, UMAPHIP_INVALID_CONFIG,
                            "adaptation samples above %lld are not supported on device",
                            (long long)kDeviceRatioMax);
            P->sample_count = S;
            for (int t = 0; t < params->epochs; ++t)
                if (t >= 10 && t <= params->epochs / 2 && t % 10 == 0)
                    P->adaptation_epochs.push_back(t);
            std::vector<int32_t> draws(P->adaptation_epochs.size() * 11 * S);
            StdRng rng(42);
            const uint32_t nn = static_cast<uint32_t>(n), qq = static_cast<uint32_t>(Q);
            for (size_t k = 0; k < P->adaptation_epochs.size(); ++k) {
                int32_t *b = draws.data() + k * 11 * S;
                for (int64_t s = 0; s < S; ++s) b[s] = static_cast<int32_t>(rng.below(nn));
                for (int64_t s = 0; s < 5 * S; ++s) b[S + s] = static_cast<int32_t>(rng.below(qq));
                for (int64_t s = 0; s < 5 * S; ++s) b[6 * S + s] = static_cast<int32_t>(rng.below(nn));
            }
            P->samples_dev.alloc(std::max<size_t>(draws.size(), 1));
            if (!draws.empty())
                HIP_CHECK(hipMemcpy(P->samples_dev.ptr, draws.data(), draws.size() * sizeof(int32_t),
                                    hipMemcpyHostToDevice));
            P->resultants.alloc(2 * S);
        }
        HIP_CHECK(hipDeviceSynchronize());
        *out = P.release();
    });
}

umaphip_status umaphip_plan_run(umaphip_plan *plan) {
    return guarded([&] {
        UMAPHIP_REQUIRE(plan != nullptr, UMAPHIP_INVALID_INPUT, "plan is null");
        plan->run();
    });
}

void *umaphip_plan_stream(umaphip_plan *plan) { return plan ? plan->stream : nullptr; }

umaphip_status umaphip_plan_set_objective(umaphip_plan *plan, const float *weights,
                                          int32_t every) {
    return guarded([&] {
        UMAPHIP_REQUIRE(plan != nullptr, UMAPHIP_INVALID_INPUT, "plan is null");
        UMAPHIP_REQUIRE(every >= 0, UMAPHIP_INVALID_CONFIG, "objective interval must be >= 0");
        plan->obj_every = every;
        if (every == 0) return;
        UMAPHIP_REQUIRE(weights != nullptr || plan->E == 0, UMAPHIP_INVALID_INPUT,
                        "objective weights are null");
        plan->obj_weights.alloc(std::max<int64_t>(plan->E, 1));
        if (plan->E)
            HIP_CHECK(hipMemcpy(plan->obj_weights.ptr, weights, plan->E * sizeof(float),
                                hipMemcpyHostToDevice));
        plan->obj_ws.alloc(cross_entropy_workspace(plan->h));
        const int64_t slots = plan->p.epochs / every + 1;
        plan->obj_vals.alloc(std::max<int64_t>(slots, 1));
    });
}

umaphip_status umaphip_plan_objective(umaphip_plan *plan, double *values, int64_t capacity,
                                      int64_t *count) {
    return guarded([&] {
        UMAPHIP_REQUIRE(plan != nullptr && count != nullptr, UMAPHIP_INVALID_INPUT,
                        "plan objective arguments are null");
        HIP_CHECK(hipStreamSynchronize(plan->stream));
        *count = plan->obj_count;
        if (values && plan->obj_count)
            HIP_CHECK(hipMemcpy(values, plan->obj_vals.ptr,
                                std::min(capacity, plan->obj_count) * sizeof(double),
                                hipMemcpyDeviceToHost));
    });
}

umaphip_status umaphip_cross_entropy_workspace_bytes(int64_t h, size_t *bytes) {
    return guarded([&] {
        UMAPHIP_REQUIRE(bytes && h >= 0, UMAPHIP_INVALID_INPUT, "invalid workspace query");
        *bytes = static_cast<size_t>(cross_entropy_workspace(h)) * sizeof(double);
    });
}

umaphip_status umaphip_cross_entropy(const float *head, int64_t h, const float *tail, int64_t n,
                                     int64_t d, const int32_t *offsets, const int32_t *indices,
                                     const float *weights, float a, float b, void *workspace,
                                     size_t workspace_bytes, double *out, void *stream) {
    return guarded([&] {
        UMAPHIP_REQUIRE(head && offsets && out && workspace && h >= 0 && d > 0,
                        UMAPHIP_INVALID_INPUT, "invalid objective arguments");
        UMAPHIP_REQUIRE(workspace_bytes >= static_cast<size_t>(cross_entropy_workspace(h)) *
                                               sizeof(double),
                        UMAPHIP_CAPACITY, "objective workspace too small");
        (void)n;
        launch_cross_entropy(head, h, tail ? tail : head, d, offsets, indices, weights, a, b,
                             static_cast<double *>(workspace), out, as_stream(stream));
    });
}

umaphip_status umaphip_plan_set_record_format(umaphip_plan *plan, int32_t compact) {
    return guarded([&] {
        UMAPHIP_REQUIRE(plan != nullptr, UMAPHIP_INVALID_INPUT, "plan is null");
        UMAPHIP_REQUIRE(plan->unit_path, UMAPHIP_INVALID_CONFIG,
                        "edge records exist only in plans that use the edge-balanced sweep");
        const bool want = compact != 0;
        UMAPHIP_REQUIRE(!want || plan->compactable, UMAPHIP_INVALID_CONFIG,
                        "24-byte records need a rank-ordered plan whose periods are rebuilt "
                        "exactly from f32 sample counts");
        if (want == plan->compact) return;
        HIP_CHECK(hipStreamSynchronize(plan->stream));
        plan->compact = want;
        plan->rec.alloc(static_cast<size_t>((want ? 3 : 4) * plan->E));
    });
}

umaphip_status umaphip_plan_set_clock_schedule(umaphip_plan *plan, int32_t span) {
    return guarded([&] {
        UMAPHIP_REQUIRE(plan != nullptr, UMAPHIP_INVALID_INPUT, "plan is null");
        UMAPHIP_REQUIRE(span <= 0 || plan->schedulable, UMAPHIP_INVALID_CONFIG,
                        "a clock schedule needs a plan on the edge-balanced sweep with "
                        "negative_sample_rate <= 64");
        HIP_CHECK(hipStreamSynchronize(plan->stream));
        plan->configure_schedule(span);
    });
}

umaphip_status umaphip_plan_set_trace(umaphip_plan *plan, int32_t epoch, int32_t negative_cap) {
    return guarded([&] {
        UMAPHIP_REQUIRE(plan != nullptr, UMAPHIP_INVALID_INPUT, "plan is null");
        if (epoch < 0) {
            plan->trace_epoch = -1;
            return;
        }
        UMAPHIP_REQUIRE(plan->unit_path, UMAPHIP_INVALID_CONFIG,
                        "the epoch trace covers the edge-balanced sweep (Adam, d = 2, more than "
                        "1024 rows)");
        UMAPHIP_REQUIRE(epoch < plan->p.epochs, UMAPHIP_INVALID_CONFIG,
                        "trace epoch %d is not below the %d epochs", epoch, plan->p.epochs);
        UMAPHIP_REQUIRE(negative_cap >= 1 && negative_cap <= 1024, UMAPHIP_INVALID_CONFIG,
                        "negative capacity must be in 1..=1024");
        const int64_t h = plan->h, n = plan->n, E = plan->E;
        const size_t slots = static_cast<size_t>(E) * static_cast<size_t>(negative_cap);
        HIP_CHECK(hipStreamSynchronize(plan->stream));
        plan->tr_head.alloc(static_cast<size_t>(2 * h));
        plan->tr_grad.alloc(static_cast<size_t>(2 * h));
        plan->tr_scale.alloc(1);
        plan->tr_order.alloc(static_cast<size_t>(n));
        plan->tr_ranks.alloc(static_cast<size_t>(h));
        plan->tr_count.alloc(static_cast<size_t>(E));
        plan->tr_cand.alloc(slots);
        plan->tr_kept.alloc(slots);
        plan->tr_misc.alloc(3);
        plan->trace_epoch = epoch;
        plan->trace_cap = negative_cap;
    });
}

umaphip_status umaphip_plan_trace(umaphip_plan *plan, float *head_before, int32_t *order,
                                  int32_t *ranks, float *scale, float *gradient,
                                  int32_t *negative_count, int32_t *negative_candidates,
                                  uint8_t *negative_kept, int32_t *info) {
    return guarded([&] {
        UMAPHIP_REQUIRE(plan != nullptr && info != nullptr, UMAPHIP_INVALID_INPUT,
                        "plan trace arguments are null");
        HIP_CHECK(hipStreamSynchronize(plan->stream));
        UMAPHIP_REQUIRE(plan->traced, UMAPHIP_INVALID_CONFIG,
                        "no traced epoch: set a trace epoch, then run the plan");
        int32_t misc[3] = {0, 0, 0};
        HIP_CHECK(hipMemcpy(misc, plan->tr_misc.ptr, sizeof(misc), hipMemcpyDeviceToHost));
        UMAPHIP_REQUIRE(misc[0] == 0, UMAPHIP_CAPACITY,
                        "an edge drew more than %d negatives in the traced epoch", plan->trace_cap);
        info[0] = plan->trace_info[0];
        info[1] = plan->trace_info[1];
        info[2] = plan->trace_info[2];
        info[3] = misc[1];
        info[4] = misc[2];
        auto down = [](void *dst, const void *src, size_t bytes) {
            if (dst && bytes) HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        };
        down(head_before, plan->tr_head.ptr, plan->tr_head.bytes());
        down(order, plan->tr_order.ptr, plan->tr_order.bytes());
        down(ranks, plan->tr_ranks.ptr, plan->tr_ranks.bytes());
        down(scale, plan->tr_scale.ptr, sizeof(float));
        down(gradient, plan->tr_grad.ptr, plan->tr_grad.bytes());
        down(negative_count, plan->tr_count.ptr, plan->tr_count.bytes());
        down(negative_candidates, plan->tr_cand.ptr, plan->tr_cand.bytes());
        down(negative_kept, plan->tr_kept.ptr, plan->tr_kept.bytes());
    });
}

umaphip_status umaphip_plan_set_profiling(umaphip_plan *plan, int32_t enabled) {
    return guarded([&] {
        UMAPHIP_REQUIRE(plan != nullptr, UMAPHIP_INVALID_INPUT, "plan is null");
        plan->profiling = enabled != 0;
    });
}

umaphip_status umaphip_plan_download(umaphip_plan *plan, float *head, double *next_positive,
                                     double *next_negative, umaphip_layout_stats *stats) {
    return guarded([&] {
        UMAPHIP_REQUIRE(plan != nullptr, UMAPHIP_INVALID_INPUT, "plan is null");
        HIP_CHECK(hipStreamSynchronize(plan->stream));
        const float *src = plan->ran ? plan->pos[plan->cur].ptr : plan->head0.ptr;
        unsigned long long status = ~0ull;
        int32_t err = 0;
        float scale = plan->scale0;
        if (plan->ran) {
            HIP_CHECK(hipMemcpy(&status, plan->status.ptr, sizeof(status), hipMemcpyDeviceToHost));
            HIP_CHECK(hipMemcpy(&err, plan->error_flag.ptr, sizeof(err), hipMemcpyDeviceToHost));
            HIP_CHECK(hipMemcpy(&scale, plan->scale_dev.ptr, sizeof(float), hipMemcpyDeviceToHost));
        }
        if (head)
            HIP_CHECK(hipMemcpy(head, src, plan->h * plan->d * sizeof(float), hipMemcpyDeviceToHost));
        if (next_positive && plan->E)
            HIP_CHECK(hipMemcpy(next_positive, plan->ran ? plan->npos.ptr : plan->npos0.ptr,
                                plan->E * sizeof(double), hipMemcpyDeviceToHost));
        if (next_negative && plan->E)
            HIP_CHECK(hipMemcpy(next_negative, plan->ran ? plan->nneg.ptr : plan->nneg0.ptr,
                                plan->E * sizeof(double), hipMemcpyDeviceToHost));
        if (stats) {
            stats->epochs_run = plan->ran ? plan->p.epochs : 0;
            stats->adaptations = static_cast<int64_t>(plan->adaptation_epochs.size());
            stats->refreshes = plan->refresh_count;
            stats->final_scale = scale;
            float ms = 0.0f;
            if (plan->ran) HIP_CHECK(hipEventElapsedTime(&ms, plan->ev_start, plan->ev_stop));
            stats->device_ms = ms;
            stats->sweep_ms = plan->ran ? plan->pairs_ms(plan->sweep_ev, plan->sweep_used) : 0.0;
            stats->sweep_launches = plan->sweep_used / 2;
            stats->refresh_ms = plan->ran ? plan->pairs_ms(plan->refresh_ev, plan->refresh_used) : 0.0;
            stats->refresh_launches = plan->refresh_used / 2;
            stats->edge_balanced = plan->unit_path ? 1 : 0;
            stats->rank_ordered_epochs = plan->ran ? plan->ranked_epochs : 0;
            stats->compact_records = plan->compact ? 1 : 0;
            stats->clock_schedule = plan->sched_span;
        }
        UMAPHIP_REQUIRE(status == ~0ull, UMAPHIP_NUMERICAL,
                        "non-finite %s projection at row %llu (refresh %llu)",
                        plan->shared ? "reference" : "reference or query",
                        status & 0xFFFFFFFFull, status >> 32);
        UMAPHIP_REQUIRE((err & 2) == 0, UMAPHIP_BACKEND, "projection sort: a pass stalled");
        UMAPHIP_REQUIRE((err & 4) == 0, UMAPHIP_BACKEND,
                        "clock schedule: an edge drew more than 254 negatives in one epoch");
        UMAPHIP_REQUIRE(err == 0, UMAPHIP_NUMERICAL, "invalid negative force ratio");
    });
}

void umaphip_plan_destroy(umaphip_plan *plan) { delete plan; }

umaphip_status umaphip_optimize_compact(int64_t h, int64_t n, int64_t d, float *head,
                                        const float *tail, const uint32_t *offsets,
                                        const uint32_t *indices, int64_t E,
                                        const double *periods, double *next_positive,
                                        const double *negative_periods, double *next_negative,
                                        const umaphip_layout_params *params, int32_t shared,
                                        umaphip_layout_stats *stats) {
    if (params && params->epochs == 0) return UMAPHIP_OK;
    umaphip_plan *plan = nullptr;
    umaphip_status s = umaphip_plan_create(h, n, d, head, tail, offsets, indices, E, periods,
                                           next_positive, negative_periods, next_negative, params,
                                           shared, &plan);
    if (s != UMAPHIP_OK) return s;
    s = umaphip_plan_run(plan);
    if (s == UMAPHIP_OK) s = umaphip_plan_download(plan, head, next_positive, next_negative, stats);
    umaphip_plan_destroy(plan);
    return s;
}

umaphip_status umaphip_orders_at_epoch(int64_t h, int64_t n, const umaphip_layout_params *params,
                                       int32_t shared, int32_t epoch, int32_t *source_out,
                                       int32_t *negative_out) {
    return guarded([&] {
        if (shared) n = h;
        check_layout_params(params, h, n, 1, 0);
        const bool momentum = params->optimizer == 1;
        const bool hard = params->negative_selection_range < n;
        const bool alias = shared && momentum && !hard;
        // Full replay (source orders always materialised) so callers see the
        // reference's orders even where the driver skips Adam's source swaps.
        OrderReplay o(h, n, params->seed, alias, hard, true);
        o.negative_live = !alias && !hard;
        for (int32_t t = 0; t < epoch; ++t) o.advance();
        if (source_out) std::memcpy(source_out, o.source.data(), sizeof(int32_t) * h);
        if (negative_out) {
            if (hard && (epoch > 0 || h > 1024)) {
                for (int64_t i = 0; i < n; ++i) negative_out[i] = static_cast<int32_t>(i);
            } else {
                std::memcpy(negative_out, o.neg().data(), sizeof(int32_t) * n);
            }
        }
    });
}

// ====================================================== epoch boundary
static void check_epoch_args(const umaphip_epoch_args *a) {
    UMAPHIP_REQUIRE(a != nullptr, UMAPHIP_INVALID_INPUT, "epoch args are null");
    UMAPHIP_REQUIRE(a->h > 0, UMAPHIP_INVALID_INPUT, "modern layout requires source rows");
    UMAPHIP_REQUIRE(a->n > 0 && a->d > 0, UMAPHIP_INVALID_INPUT, "invalid reference/component count");
    UMAPHIP_REQUIRE(a->negative_range >= 1 && a->negative_range <= a->n, UMAPHIP_INVALID_INPUT,
                    "negative range must be in 1..=N");
    UMAPHIP_REQUIRE(a->head && a->offsets && a->source_order && a->negative_order && a->updates,
                    UMAPHIP_INVALID_INPUT, "required epoch pointer is null");
    UMAPHIP_REQUIRE(a->shared || a->fixed_tail, UMAPHIP_INVALID_INPUT, "transform requires a tail");
    UMAPHIP_REQUIRE(a->momentum || (a->m && a->v), UMAPHIP_INVALID_INPUT, "adam requires m and v");
    UMAPHIP_REQUIRE(!(a->h > 1024 && a->negative_range < a->n) || a->source_ranks,
                    UMAPHIP_INVALID_INPUT, "source ranks must cover every source row");
}

static RowwiseArgs rowwise_from(const umaphip_epoch_args *a) {
    RowwiseArgs R{};
    R.h = a->h; R.n = a->n; R.d = a->d;
    R.head = a->head; R.hs0 = a->head_s0; R.hs1 = a->head_s1;
    R.tail = a->shared ? a->head : a->fixed_tail;
    R.ts0 = a->shared ? a->head_s0 : a->tail_s0;
    R.ts1 = a->shared ? a->head_s1 : a->tail_s1;
    R.offsets = a->offsets; R.indices = a->indices;
    R.source_order = a->source_order; R.negative_order = a->negative_order;
    R.source_ranks = a->source_ranks;
    R.periods = a->periods; R.next_positive = a->next_positive;
    R.negative_periods = a->negative_periods; R.next_negative = a->next_negative;
    R.updates = a->updates; R.us0 = a->upd_s0; R.us1 = a->upd_s1;
    R.epoch = a->epoch; R.a = a->a; R.b = a->b; R.gamma = a->gamma; R.scale = a->scale;
    R.range = a->negative_range;
    R.shared = a->shared; R.negatives = a->negatives; R.exclude = a->exclude_graph_neighbors;
    return R;
}

umaphip_status umaphip_modern_epoch(const umaphip_epoch_args *a, void *stream) {
    return guarded([&] {
        check_epoch_args(a);
        hipStream_t s = as_stream(stream);
        RowwiseArgs R = rowwise_from(a);
        const int64_t h = a->h, d = a->d;
        if (a->momentum) {
            const int64_t block = h <= 1024 ? h : std::max<int64_t>(h / 2, 1);
            R.multiplier = a->alpha;
            for (int64_t start = 0; start < h; start += block) {
                const int64_t end = std::min(start + block, h);
                launch_gradient_rowwise(R, start, end, s);
                launch_momentum_apply(d, a->head, a->head_s0, a->head_s1, a->updates, a->upd_s0,
                                      a->upd_s1, a->source_order, start, end, s);
            }
            launch_momentum_decay(h, d, a->updates, a->upd_s0, a->upd_s1, a->mu, s);
            return;
        }
        const bool contiguous2 = d == 2 && a->head_s0 == 2 && a->head_s1 == 1 &&
                                 a->upd_s0 == 2 && a->upd_s1 == 1 &&
                                 (a->shared || (a->tail_s0 == 2 && a->tail_s1 == 1));
        if (contiguous2) {
            // Adam is whole-epoch Jacobi: the gradient does not depend on the
            // source order, so rows are swept in CSR order.
            SweepArgs A{};
            A.h = h; A.n = a->n;
            A.pos = a->head; A.tgt = a->shared ? a->head : a->fixed_tail;
            A.order = a->negative_order; A.ranks = a->source_ranks;
            A.offsets = a->offsets; A.indices = a->indices;
            A.periods = a->periods; A.next_positive = a->next_positive;
            A.negative_periods = a->negative_periods; A.next_negative = a->next_negative;
            A.updates = a->updates; A.scale = a->scale;
            A.epoch = a->epoch; A.a = a->a; A.b = a->b; A.gamma = a->gamma;
            A.range = a->negative_range; A.multiplier = 1.0f;
            A.shared = a->shared; A.negatives = a->negatives;
            A.exclude = a->exclude_graph_neighbors; A.fused = 0;
            launch_edge_sweep(A, s);
        } else {
            R.multiplier = 1.0f;
            launch_gradient_rowwise(R, 0, h, s);
        }
        launch_adam(h, d, a->head, a->head_s0, a->head_s1, a->updates, a->upd_s0, a->upd_s1, a->m,
                    a->m_s0, a->m_s1, a->v, a->v_s0, a->v_s1, a->alpha, a->correction1,
                    a->correction2, s);
    });
}

umaphip_status umaphip_debug_negatives(const umaphip_epoch_args *a, int32_t cap,
                                       int32_t *count_out, int32_t *cand_out, uint8_t *kept_out,
                                       void *stream) {
    return guarded([&] {
        check_epoch_args(a);
        UMAPHIP_REQUIRE(cap > 0 && count_out && cand_out && kept_out, UMAPHIP_INVALID_INPUT,
                        "debug outputs are null");
        hipStream_t s = as_stream(stream);
        DeviceArray<int32_t> overflow(1);
        HIP_CHECK(hipMemsetAsync(overflow.ptr, 0, sizeof(int32_t), s));
        launch_debug_negatives(rowwise_from(a), cap, count_out, cand_out, kept_out, overflow.ptr, s);
        int32_t host = 0;
        HIP_CHECK(hipMemcpyAsync(&host, overflow.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        UMAPHIP_REQUIRE(host == 0, UMAPHIP_CAPACITY, "an edge visits more than %d candidates", cap);
    });
}

umaphip_status umaphip_project(const float *x, int64_t rows, int64_t d, int64_t s0, int64_t s1,
                               const float *direction, float *out, void *stream) {
    return guarded([&] {
        UMAPHIP_REQUIRE(rows >= 0 && d > 0, UMAPHIP_INVALID_INPUT, "invalid projection shape");
        launch_project(x, rows, d, s0, s1, direction, out, as_stream(stream));
    });
}

umaphip_status umaphip_argsort_workspace_bytes(int64_t n, size_t *bytes) {
    return guarded([&] {
        UMAPHIP_REQUIRE(n >= 0 && n <= INT32_MAX && bytes, UMAPHIP_INVALID_INPUT,
                        "invalid argsort size");
        // keys, sorted keys, iota, then the rocPRIM scratch
        const size_t align = 256;
        auto up = [&](size_t x) { return (x + align - 1) / align * align; };
        *bytes = 3 * up(static_cast<size_t>(n) * 4) + up(sort_pairs_workspace(n));
    });
}

umaphip_status umaphip_stable_argsort_f32(const float *values, int64_t n, int32_t *order,
                                          float *sorted, int64_t *status_out, void *workspace,
                                          size_t workspace_bytes, void *stream) {
    return guarded([&] {
        size_t need = 0;
        if (umaphip_argsort_workspace_bytes(n, &need) != UMAPHIP_OK) throw Fail{UMAPHIP_INVALID_INPUT};
        UMAPHIP_REQUIRE(workspace_bytes >= need && workspace, UMAPHIP_CAPACITY,
                        "argsort workspace too small (%zu < %zu)", workspace_bytes, need);
        hipStream_t s = as_stream(stream);
        const size_t align = 256;
        auto up = [&](size_t x) { return (x + align - 1) / align * align; };
        char *w = static_cast<char *>(workspace);
        uint32_t *keys = reinterpret_cast<uint32_t *>(w);
        uint32_t *keys_sorted = reinterpret_cast<uint32_t *>(w + up(n * 4));
        int32_t *iota = reinterpret_cast<int32_t *>(w + 2 * up(n * 4));
        void *scratch = w + 3 * up(n * 4);
        unsigned long long *st = reinterpret_cast<unsigned long long *>(status_out);
        launch_status_min_u64(st, ~0ull, s);
        launch_keys_from_values(values, n, keys, st, s);
        launch_iota(iota, n, s);
        launch_sort_pairs(scratch, workspace_bytes - 3 * up(n * 4), keys, keys_sorted, iota, order,
                          n, s);
        launch_values_from_keys(keys_sorted, n, sorted, s);
    });
}

}  // extern "C"

// Scratch of the allocation-free argsort: the workspace of
// umaphip_stable_argsort_f32 for up to kNoAllocSortKeys keys, as a
// file-scope device array that every device holds until the process ends.
// Host code never uses its name as an address: each device's copy is
// resolved once with hipGetSymbolAddress and kept per device.
namespace {
constexpr int64_t kNoAllocSortKeys = int64_t(1) << 20;
constexpr size_t kNoAllocAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kNoAllocAlign - 1) / kNoAllocAlign * kNoAllocAlign; }
// argsort workspace: keys, sorted keys, iota, then the radix sort's own
// (ping-pong keys and values, per-tile digit histograms, digit totals)
constexpr size_t kSortTileKeys = 2048;
constexpr size_t kNoAllocScratchBytes =
    3 * align_up(kNoAllocSortKeys * 4) + 2 * align_up(kNoAllocSortKeys * 4) +
    align_up(256 * ((kNoAllocSortKeys + kSortTileKeys - 1) / kSortTileKeys) * 4) +
    align_up(1024 * 4);
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
// This device's copy of `scratch` (resolved on the first call on a device,
// with that device current).
void *scratch_for_current_device() {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    UMAPHIP_REQUIRE(dev >= 0 && dev < kMaxDevices, UMAPHIP_BACKEND, "device id %d out of range",
                    dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        UMAPHIP_REQUIRE(rc == hipSuccess && ptr != nullptr, UMAPHIP_BACKEND,
                        "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(rc));
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

umaphip_status umaphip_stable_argsort_f32_noalloc(const float *values, int64_t n, int32_t *order,
                                                  float *sorted, int64_t *status_out,
                                                  void *stream) {
    return guarded([&] {
        UMAPHIP_REQUIRE(n >= 0, UMAPHIP_INVALID_INPUT, "invalid argsort size");
        UMAPHIP_REQUIRE(n <= kNoAllocSortKeys, UMAPHIP_CAPACITY,
                        "the allocation-free argsort holds at most %lld keys (got %lld); use "
                        "umaphip_stable_argsort_f32 with a workspace",
                        (long long)kNoAllocSortKeys, (long long)n);
        size_t need = 0;
        if (umaphip_argsort_workspace_bytes(n, &need) != UMAPHIP_OK) throw Fail{UMAPHIP_INVALID_INPUT};
        UMAPHIP_REQUIRE(need <= kNoAllocScratchBytes, UMAPHIP_BACKEND,
                        "argsort scratch too small (%zu < %zu)", kNoAllocScratchBytes, need);
        void *ws = scratch_for_current_device();
        const umaphip_status st = umaphip_stable_argsort_f32(values, n, order, sorted, status_out,
                                                             ws, kNoAllocScratchBytes, stream);
        if (st != UMAPHIP_OK) throw Fail{st};
    });
}

umaphip_status umaphip_projection_ranks(const int32_t *order, const float *sorted, int64_t n,
                                        const float *query, int64_t h, int32_t *ranks,
                                        int32_t shared, void *stream) {
    return guarded([&] {
        UMAPHIP_REQUIRE(n > 0, UMAPHIP_INVALID_INPUT, "ranks require reference rows");
        if (shared) launch_inverse_ranks(order, n, ranks, as_stream(stream));
        else launch_lower_bound(sorted, n, query, h, ranks, as_stream(stream));
    });
}

umaphip_status umaphip_sample_resultants(const float *embedding, int64_t n, int64_t d,
                                         const int32_t *order, const int32_t *offsets,
                                         const int32_t *indices, const int32_t *sample_ranks,
                                         const int32_t *local_offsets,
                                         const int32_t *global_ranks, int64_t samples,
                                         uint32_t raw_range, float a, float b, float gamma,
                                         int32_t exclude, double *resultants, void *stream) {
    return guarded([&] {
        UMAPHIP_REQUIRE(raw_range >= 1 && raw_range <= n, UMAPHIP_INVALID_INPUT,
                        "raw range must be in 1..=N");
        launch_resultants(embedding, n, d, order, offsets, indices, sample_ranks, local_offsets,
                          global_ranks, samples, raw_range, a, b, gamma, exclude, resultants,
                          as_stream(stream));
    });
}

}  // extern "C"
