// Batched layouts: many independent fit layouts (graphs of uneven size) run
// as one launch per epoch, and the whole epoch loop is captured once as a
// hipGraph and replayed.
//
// Each graph is a modern Adam fit in global negative mode (its rows do not
// exceed negative_selection_range), so everything that is not a position is
// known before the first epoch: the per-epoch negative orders come from the
// graph's own StdRng(seed) shuffles (modern_orders.rs:359-680) and are
// generated on host threads up front, and each epoch's arguments (schedule
// value, bias corrections, order, Jacobi buffers) are fixed. In each epoch
// every graph is a segment of tiles; a graph's rows are computed exactly as
// a launch of that graph alone would compute them. Two launch forms give
// the same bits: one segment-sweep launch per epoch (replayed from a captured
// hipGraph), or one persistent cooperative launch that runs every epoch with
// a grid barrier between epochs. The cooperative launch is issued directly,
// never captured: a graph node replayed without the cooperative launch's
// set-up would have no grid barrier to meet at. The tile shape (rows per
// tile x lanes) is chosen by timing the candidate shapes on the first epochs
// in the chosen form (umaphip_batch_tune).
//
// The clocks do not depend on the positions: when every epoch's term counts
// fit kBatchScheduleBytes (one byte per edge and epoch), a schedule pass
// before the first epoch runs all clocks ahead (launch_edge_schedule) and
// the tile sweeps read each edge's count instead of its four f64 clock
// words. Results are the same bits either way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "common.hpp"
#include "layout_host.hpp"
#include "layout_kernels.hpp"
#include "stdrng.hpp"

using namespace umaphip;

namespace {
constexpr int64_t kBatchScheduleBytes = int64_t(8) << 30;

// Whether every term count the clocks can produce fits a byte: in steady
// state a firing edge draws at most (period + 1) / negative_period + 1
// negatives, and the first firing at most (next_positive + 1 -
// next_negative) / negative_period + 1.
bool counts_fit_byte(const double *per, const double *nper, const double *npos,
                     const double *nneg, int64_t E) {
    for (int64_t e = 0; e < E; ++e) {
        if (!(per[e] > 0.0) || !(nper[e] > 0.0)) return false;
        const double steady = (per[e] + 1.0) / nper[e] + 1.0;
        const double first = (npos[e] + 1.0 - nneg[e]) / nper[e] + 1.0;
        if (!(steady < 250.0) || !(first < 250.0)) return false;
    }
    return true;
}
}  // namespace

struct umaphip_batch {
    int32_t G = 0;
    int64_t d = 2, total_rows = 0, total_edges = 0, total_order = 0;
    int epochs = 0;
    umaphip_layout_params p{};
    std::vector<int64_t> rows, edges, row_start, edge_start, off_start, order_start;
    std::vector<float> scale;
    std::vector<uint32_t> range;
    std::vector<float> alpha, mu, c1, c2;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DeviceArray<float> pos[2], m, v, head0;
    DeviceArray<int32_t> offsets, indices, orders;
    DeviceArray<double> periods, negperiods, npos, nneg, npos0, nneg0;
    DeviceArray<SweepSegment> segs;
    DeviceArray<int32_t> counters;      // persistent launch: tile counters per epoch and range
    // clock schedule: every epoch's term counts, computed by one pass before
    // the first epoch (sched_span = epochs), or off (0)
    bool schedulable = false;
    int32_t sched_span = 0;
    int64_t sched_stride = 0;
    DeviceArray<uint8_t> sched;
    DeviceArray<int32_t> flag;          // bit 4: a term count overflowed its byte
    // per-epoch negative tables (table[i] = position at rank i - r0 of the
    // graph's order) for the per-epoch launch form: one gather per negative
    DeviceArray<float> table;
    DeviceArray<int32_t> graph_of_row;
    DeviceArray<int64_t> row_start_dev, order_start_dev;
    bool built_persistent = false;
    int shape = 0;
    int64_t tiles = 0;
    bool built = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool persistent = false;            // one cooperative launch for every epoch
    float last_ms = 0.0f;
    bool ran = false;

    ~umaphip_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Segment arguments of every epoch for a tile shape.
    void build(int s) {
        if (built && shape == s && built_persistent == persistent) return;
        if (exec) {
            HIP_CHECK(hipGraphExecDestroy(exec));
            exec = nullptr;
        }
        if (graph) {
            HIP_CHECK(hipGraphDestroy(graph));
            graph = nullptr;
        }
        shape = s;
        built_persistent = persistent;
        const int64_t tr = tile_rows(s);
        std::vector<SweepSegment> host(static_cast<size_t>(epochs) * G);
        tiles = 0;
        std::vector<int64_t> tile0(G);
        for (int32_t g = 0; g < G; ++g) {
            tile0[g] = tiles;
            tiles += (rows[g] + tr - 1) / tr;
        }
        for (int t = 0; t < epochs; ++t) {
            for (int32_t g = 0; g < G; ++g) {
                SweepArgs A{};
                const int64_t r0 = row_start[g], e0 = edge_start[g];
                A.h = rows[g];
                A.n = rows[g];
                A.pos = pos[t & 1].ptr + r0 * d;
                A.pos_out = pos[(t + 1) & 1].ptr + r0 * d;
                A.tgt = A.pos;
                // the persistent launch reads positions through the order
                A.table = persistent ? nullptr : table.ptr + r0 * d;
                A.order = orders.ptr + order_start[g] + static_cast<int64_t>(t) * rows[g];
                A.ranks = nullptr;
                A.offsets = offsets.ptr + off_start[g];
                A.indices = indices.ptr + e0;
                A.periods = periods.ptr + e0;
                A.next_positive = npos.ptr + e0;
                A.negative_periods = negperiods.ptr + e0;
                A.next_negative = nneg.ptr + e0;
                A.m = m.ptr + r0 * d;
                A.v = v.ptr + r0 * d;
                A.scale = scale[g];
                A.epoch = static_cast<uint32_t>(t);
                A.a = p.a; A.b = p.b; A.gamma = p.repulsion; A.range = range[g];
                A.multiplier = 1.0f; A.alpha = alpha[t]; A.c1 = c1[t]; A.c2 = c2[t];
                A.shared = 1; A.negatives = p.negative_sample_rate != 0;
                A.exclude = p.exclude_graph_neighbors; A.fused = 1;
                A.sched = sched_span ? sched.ptr + (t % sched_span) * sched_stride + e0 : nullptr;
                host[static_cast<size_t>(t) * G + g].args = A;
                host[static_cast<size_t>(t) * G + g].tile0 = tile0[g];
            }
        }
        segs.alloc(std::max<size_t>(host.size(), 1));
        counters.alloc(static_cast<size_t>(std::max(epochs, 1)) * 8);
        if (!host.empty())
            HIP_CHECK(hipMemcpy(segs.ptr, host.data(), host.size() * sizeof(SweepSegment),
                                hipMemcpyHostToDevice));
        built = true;
    }

    // Schedule on (when the batch allows it) or off; segments are rebuilt.
    void set_schedule(bool on) {
        const int32_t span = on && schedulable ? epochs : 0;
        if (span == sched_span && (span == 0 || sched.ptr)) return;
        HIP_CHECK(hipStreamSynchronize(stream));
        sched_span = span;
        sched_stride = span ? unit_schedule_stride(total_edges) : 0;
        try {
            sched.alloc(static_cast<size_t>(sched_stride * span));
        } catch (const Fail &f) {
            // an optimisation: without the memory the sweeps advance the clocks
            sched_span = 0;
            sched_stride = 0;
            if (f.status != UMAPHIP_MEMORY) throw;
            (void)hipGetLastError();
            set_error("%s", "");
        }
        if (exec) {
            HIP_CHECK(hipGraphExecDestroy(exec));
            exec = nullptr;
        }
        if (graph) {
            HIP_CHECK(hipGraphDestroy(graph));
            graph = nullptr;
        }
        built = false;
    }

    // the schedule pass that epoch t starts, if any
    void schedule_from(int t, int count) {
        if (sched_span && t % sched_span == 0)
            launch_edge_schedule(npos.ptr, nneg.ptr, periods.ptr, negperiods.ptr, total_edges,
                                 p.negative_sample_rate != 0, t, std::min(sched_span, count - t),
                                 sched.ptr, sched_stride, flag.ptr, stream);
    }

    void reset() {
        HIP_CHECK(hipMemsetAsync(flag.ptr, 0, sizeof(int32_t), stream));
        HIP_CHECK(hipMemcpyAsync(pos[0].ptr, head0.ptr, head0.bytes(), hipMemcpyDeviceToDevice, stream));
        HIP_CHECK(hipMemsetAsync(m.ptr, 0, m.bytes(), stream));
        HIP_CHECK(hipMemsetAsync(v.ptr, 0, v.bytes(), stream));
        if (total_edges) {
            HIP_CHECK(hipMemcpyAsync(npos.ptr, npos0.ptr, npos0.bytes(), hipMemcpyDeviceToDevice, stream));
            HIP_CHECK(hipMemcpyAsync(nneg.ptr, nneg0.ptr, nneg0.bytes(), hipMemcpyDeviceToDevice, stream));
        }
    }

    void enqueue_epochs(int count) {
        if (persistent) {
            // sched_span is 0 or every epoch: one pass covers the launch
            schedule_from(0, count);
            HIP_CHECK(launch_persistent_segments(segs.ptr, G, tiles, count, counters.ptr, stream,
                                                 shape));
            return;
        }
        for (int t = 0; t < count; ++t) {
            schedule_from(t, count);
            launch_batch_table(pos[t & 1].ptr, orders.ptr, row_start_dev.ptr, order_start_dev.ptr,
                               graph_of_row.ptr, total_rows, t, table.ptr, stream);
            launch_segment_sweep(segs.ptr + static_cast<size_t>(t) * G, G, tiles, stream, shape);
        }
    }

    // the per-epoch launches of every epoch as one graph
    void capture() {
        if (exec) return;
        HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        enqueue_epochs(epochs);
        HIP_CHECK(hipStreamEndCapture(stream, &graph));
        HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }

    // persistent: one cooperative launch whatever use_graph says
    void run(bool use_graph) {
        build(shape);   // no-op unless the shape or the launch form changed
        if (use_graph && !persistent) capture();
        reset();
        HIP_CHECK(hipEventRecord(ev0, stream));
        if (use_graph && !persistent) HIP_CHECK(hipGraphLaunch(exec, stream));
        else enqueue_epochs(epochs);
        HIP_CHECK(hipEventRecord(ev1, stream));
        ran = true;
    }
};

extern "C" {

umaphip_status umaphip_batch_create(int32_t G, const int64_t *rows, const int64_t *n_edges,
                                    const uint32_t *offsets, const uint32_t *indices,
                                    const double *periods, const double *next_positive,
                                    const double *negative_periods, const double *next_negative,
                                    const float *head, int64_t d,
                                    const umaphip_layout_params *params, const uint64_t *seeds,
                                    umaphip_batch **out) {
    return guarded([&] {
        UMAPHIP_REQUIRE(out && rows && n_edges && offsets && head && params && seeds,
                        UMAPHIP_INVALID_INPUT, "batch arguments are null");
        *out = nullptr;
        UMAPHIP_REQUIRE(G >= 1, UMAPHIP_INVALID_INPUT, "a batch needs at least one graph");
        UMAPHIP_REQUIRE(d == 2, UMAPHIP_INVALID_CONFIG, "batched layouts support 2-D embeddings");
        const umaphip_layout_params &p = *params;
        UMAPHIP_REQUIRE(p.optimizer == 0, UMAPHIP_INVALID_CONFIG, "batched layouts run Adam");
        UMAPHIP_REQUIRE(p.epochs >= 0 && p.negative_selection_range > 0 &&
                            p.negative_sample_rate >= 0,
                        UMAPHIP_INVALID_CONFIG, "invalid layout parameters");
        std::unique_ptr<umaphip_batch> B(new umaphip_batch());
        B->G = G; B->d = d; B->p = p; B->epochs = p.epochs;
        B->rows.assign(rows, rows + G);
        B->edges.assign(n_edges, n_edges + G);
        B->row_start.resize(G + 1);
        B->edge_start.resize(G + 1);
        B->off_start.resize(G + 1);
        B->order_start.resize(G + 1);
        B->row_start[0] = B->edge_start[0] = B->off_start[0] = B->order_start[0] = 0;
        const uint64_t Q = static_cast<uint64_t>(p.negative_selection_range);
        for (int32_t g = 0; g < G; ++g) {
            const int64_t h = rows[g], E = n_edges[g];
            UMAPHIP_REQUIRE(h >= 1 && h <= INT32_MAX && E >= 0 && E <= INT32_MAX,
                            UMAPHIP_INVALID_INPUT, "graph %d: invalid sizes", g);
            UMAPHIP_REQUIRE(static_cast<uint64_t>(h) <= Q, UMAPHIP_INVALID_CONFIG,
                            "graph %d: %lld rows exceed negative_selection_range; batched "
                            "layouts run global negative sampling", g, (long long)h);
            const uint32_t *off = offsets + B->off_start[g];
            UMAPHIP_REQUIRE(off[0] == 0 && static_cast<int64_t>(off[h]) == E, UMAPHIP_INVALID_INPUT,
                            "graph %d: CSR offsets must run from 0 to E", g);
            for (int64_t i = 0; i < h; ++i)
                UMAPHIP_REQUIRE(off[i] <= off[i + 1], UMAPHIP_INVALID_INPUT,
                                "graph %d: CSR offsets must be nondecreasing", g);
            for (int64_t e = 0; e < E; ++e)
                UMAPHIP_REQUIRE(static_cast<int64_t>(indices[B->edge_start[g] + e]) < h,
                                UMAPHIP_INVALID_INPUT, "graph %d: CSR index out of range", g);
            B->row_start[g + 1] = B->row_start[g] + h;
            B->edge_start[g + 1] = B->edge_start[g] + E;
            B->off_start[g + 1] = B->off_start[g] + h + 1;
            B->order_start[g + 1] = B->order_start[g] + h * std::max(p.epochs, 1);
            B->range.push_back(static_cast<uint32_t>(effective_range(Q, h)));
            B->scale.push_back(resolved_scale(Q, h, p.has_scale != 0, p.scale,
                                              p.adaptation_samples > 0));
        }
        B->total_rows = B->row_start[G];
        B->total_edges = B->edge_start[G];
        B->total_order = B->order_start[G];
        schedules(p.epochs, p.initial_alpha, p.good_initialization != 0, false, B->alpha, B->mu);
        bias_corrections(p.epochs, B->c1, B->c2);

        // negative orders of every epoch, one host thread per group of graphs
        std::vector<int32_t> orders(static_cast<size_t>(std::max<int64_t>(B->total_order, 1)));
        {
            const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::vector<std::thread> pool;
            for (unsigned w = 0; w < hw; ++w)
                pool.emplace_back([&, w] {
                    for (int32_t g = static_cast<int32_t>(w); g < G; g += static_cast<int32_t>(hw)) {
                        const int64_t h = B->rows[g];
                        OrderReplay rep(h, h, seeds[g], false, false, false);
                        for (int t = 0; t < p.epochs; ++t) {
                            std::memcpy(orders.data() + B->order_start[g] + t * h, rep.neg().data(),
                                        h * sizeof(int32_t));
                            rep.advance();
                        }
                    }
                });
            for (auto &t : pool) t.join();
        }
        HIP_CHECK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreate(&B->ev0));
        HIP_CHECK(hipEventCreate(&B->ev1));
        const size_t hd = static_cast<size_t>(B->total_rows * d);
        B->head0.alloc(hd);
        B->pos[0].alloc(hd);
        B->pos[1].alloc(hd);
        B->m.alloc(hd);
        B->v.alloc(hd);
        HIP_CHECK(hipMemcpy(B->head0.ptr, head, hd * sizeof(float), hipMemcpyHostToDevice));
        B->offsets.alloc(B->off_start[G]);
        HIP_CHECK(hipMemcpy(B->offsets.ptr, offsets, B->off_start[G] * sizeof(int32_t),
                            hipMemcpyHostToDevice));
        const size_t Ec = static_cast<size_t>(std::max<int64_t>(B->total_edges, 1));
        B->indices.alloc(Ec);
        B->periods.alloc(Ec);
        B->negperiods.alloc(Ec);
        B->npos.alloc(Ec);
        B->nneg.alloc(Ec);
        B->npos0.alloc(Ec);
        B->nneg0.alloc(Ec);
        if (B->total_edges) {
            const size_t E = static_cast<size_t>(B->total_edges);
            HIP_CHECK(hipMemcpy(B->indices.ptr, indices, E * 4, hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(B->periods.ptr, periods, E * 8, hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(B->negperiods.ptr, negative_periods, E * 8, hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(B->npos0.ptr, next_positive, E * 8, hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(B->nneg0.ptr, next_negative, E * 8, hipMemcpyHostToDevice));
        }
        B->table.alloc(hd);
        {
            std::vector<int32_t> gor(static_cast<size_t>(B->total_rows));
            for (int32_t g = 0; g < G; ++g)
                std::fill(gor.begin() + B->row_start[g], gor.begin() + B->row_start[g + 1], g);
            B->graph_of_row.alloc(gor.size());
            HIP_CHECK(hipMemcpy(B->graph_of_row.ptr, gor.data(), gor.size() * sizeof(int32_t),
                                hipMemcpyHostToDevice));
            B->row_start_dev.alloc(G + 1);
            B->order_start_dev.alloc(G + 1);
            HIP_CHECK(hipMemcpy(B->row_start_dev.ptr, B->row_start.data(), (G + 1) * sizeof(int64_t),
                                hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(B->order_start_dev.ptr, B->order_start.data(),
                                (G + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
        }
        B->flag.alloc(1);
        HIP_CHECK(hipMemset(B->flag.ptr, 0, sizeof(int32_t)));
        B->schedulable = B->total_edges > 0 && p.epochs > 0 &&
                         unit_schedule_stride(B->total_edges) * p.epochs <= kBatchScheduleBytes &&
                         counts_fit_byte(periods, negative_periods, next_positive, next_negative,
                                         B->total_edges);
        B->set_schedule(true);
        B->orders.alloc(orders.size());
        HIP_CHECK(hipMemcpy(B->orders.ptr, orders.data(), orders.size() * sizeof(int32_t),
                            hipMemcpyHostToDevice));
        *out = B.release();
    });
}

umaphip_status umaphip_batch_tune(umaphip_batch *batch, int32_t probe_epochs, int32_t *shape_out,
                                  double *ms_out) {
    return guarded([&] {
        UMAPHIP_REQUIRE(batch && shape_out, UMAPHIP_INVALID_INPUT, "batch tune arguments are null");
        const int probe = std::max(1, std::min(probe_epochs, batch->epochs));
        double best = 1e300;
        int best_shape = 0;
        for (int s = 0; s < kTileShapes; ++s) {
            batch->build(s);
            batch->reset();
            batch->enqueue_epochs(probe);   // warm
            batch->reset();
            HIP_CHECK(hipEventRecord(batch->ev0, batch->stream));
            batch->enqueue_epochs(probe);
            HIP_CHECK(hipEventRecord(batch->ev1, batch->stream));
            HIP_CHECK(hipEventSynchronize(batch->ev1));
            float ms = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&ms, batch->ev0, batch->ev1));
            if (ms_out) ms_out[s] = ms;
            if (ms < best) {
                best = ms;
                best_shape = s;
            }
        }
        batch->build(best_shape);
        batch->ran = false;
        *shape_out = best_shape;
    });
}

umaphip_status umaphip_batch_set_shape(umaphip_batch *batch, int32_t shape) {
    return guarded([&] {
        UMAPHIP_REQUIRE(batch, UMAPHIP_INVALID_INPUT, "batch is null");
        UMAPHIP_REQUIRE(shape >= 0 && shape < kTileShapes, UMAPHIP_INVALID_CONFIG,
                        "tile shape must be in 0..%d", kTileShapes - 1);
        batch->build(shape);
    });
}

umaphip_status umaphip_batch_set_persistent(umaphip_batch *batch, int32_t persistent) {
    return guarded([&] {
        UMAPHIP_REQUIRE(batch, UMAPHIP_INVALID_INPUT, "batch is null");
        HIP_CHECK(hipStreamSynchronize(batch->stream));
        batch->persistent = persistent != 0;   // segments are rebuilt for the form
    });
}

umaphip_status umaphip_batch_set_clock_schedule(umaphip_batch *batch, int32_t enabled) {
    return guarded([&] {
        UMAPHIP_REQUIRE(batch, UMAPHIP_INVALID_INPUT, "batch is null");
        const int shape = batch->shape;
        batch->set_schedule(enabled != 0);
        batch->build(shape);
    });
}

int32_t umaphip_batch_clock_schedule(const umaphip_batch *batch) {
    return batch ? batch->sched_span : 0;
}

umaphip_status umaphip_batch_run(umaphip_batch *batch, int32_t use_graph) {
    return guarded([&] {
        UMAPHIP_REQUIRE(batch, UMAPHIP_INVALID_INPUT, "batch is null");
        batch->run(use_graph != 0);
    });
}

umaphip_status umaphip_batch_download(umaphip_batch *batch, float *head, double *next_positive,
                                      double *next_negative, double *device_ms) {
    return guarded([&] {
        UMAPHIP_REQUIRE(batch, UMAPHIP_INVALID_INPUT, "batch is null");
        HIP_CHECK(hipStreamSynchronize(batch->stream));
        if (device_ms) {
            float ms = 0.0f;
            if (batch->ran) HIP_CHECK(hipEventElapsedTime(&ms, batch->ev0, batch->ev1));
            *device_ms = ms;
        }
        if (batch->ran) {
            int32_t flag = 0;
            HIP_CHECK(hipMemcpy(&flag, batch->flag.ptr, sizeof(flag), hipMemcpyDeviceToHost));
            UMAPHIP_REQUIRE((flag & 4) == 0, UMAPHIP_BACKEND,
                            "clock schedule: an edge drew more than 254 negatives in one epoch");
        }
        const float *src = batch->ran ? batch->pos[batch->epochs & 1].ptr : batch->head0.ptr;
        if (head)
            HIP_CHECK(hipMemcpy(head, src, batch->total_rows * batch->d * sizeof(float),
                                hipMemcpyDeviceToHost));
        if (batch->total_edges) {
            if (next_positive)
                HIP_CHECK(hipMemcpy(next_positive, batch->ran ? batch->npos.ptr : batch->npos0.ptr,
                                    batch->total_edges * sizeof(double), hipMemcpyDeviceToHost));
            if (next_negative)
                HIP_CHECK(hipMemcpy(next_negative, batch->ran ? batch->nneg.ptr : batch->nneg0.ptr,
                                    batch->total_edges * sizeof(double), hipMemcpyDeviceToHost));
        }
    });
}

void umaphip_batch_destroy(umaphip_batch *batch) { delete batch; }

}  // extern "C"
