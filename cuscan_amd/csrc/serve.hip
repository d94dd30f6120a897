// Continuous-batching transform server over a resident fitted model.
//
// A transform request (cuscan FittedUMAP::transform, crates/umap/src/
// standard_umap/estimator.rs:1367-2264) is: exact kNN of its rows against
// the fitted references, a bipartite fuzzy graph, a weighted initial
// placement, then E_tr Adam epochs against the fixed fitted embedding. Two
// facts make requests batchable without changing any result:
//   * every query row's kNN depends only on that row, and
//   * in a transform layout every row moves alone against the fixed tail,
//     and the tail's projection order for each epoch depends only on the
//     tail and on StdRng(transform_seed) (modern_negatives.rs:283-351), not
//     on the request.
// The server therefore keeps the references (BF16 + f32), the embedding and
// the per-direction projection tables (order, sorted keys, tail positions by
// rank) resident. Requests are admitted whenever HBM in a fixed arena and a
// kNN batch slot are free; a batch's kNN runs on its own stream, then, on one
// of a few layout streams, every request's graph is built on the device from
// the kNN outputs (graph_device.hip) and the whole batch's transform layouts
// (every epoch of every request) run in one launch, so batches overlap on
// the device and the host only fills descriptor tables. A request's rows
// are computed exactly as the single-request transform computes them (same
// graph kernels, same tile sweep and per-row summation order), so results
// do not depend on what else is in flight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "graph_device.hpp"
#include "knn.hpp"
#include "layout_host.hpp"
#include "layout_kernels.hpp"
#include "stdrng.hpp"

namespace umaphip {
namespace {

double now_seconds() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// First-in first-out allocator over one device buffer. Transform requests
// all run the same number of epochs and are admitted in order, so they
// finish nearly in order; a block is released when it and every older block
// have been freed (a request that fails before its layout frees early).
struct FifoArena {
    struct Block {
        size_t off, bytes;
        bool freed;
    };
    uint8_t *base = nullptr;
    size_t cap = 0, head = 0, start = 0;
    std::deque<Block> live;
    bool alloc(size_t bytes, size_t &off) {
        bytes = align256(bytes);
        if (bytes > cap) return false;
        if (live.empty()) head = start = 0;
        if (live.empty() || head > start) {
            if (cap - head >= bytes) {
                off = head;
            } else if (start >= bytes) {
                off = 0;   // wrap; [head, cap) stays unused until start passes it
            } else {
                return false;
            }
        } else {
            if (start - head < bytes) return false;
            off = head;
        }
        head = off + bytes;
        live.push_back({off, bytes, false});
        return true;
    }
    void release(size_t off) {
        for (auto &b : live)
            if (b.off == off && !b.freed) {
                b.freed = true;
                break;
            }
        while (!live.empty() && live.front().freed) live.pop_front();
        if (live.empty()) head = start = 0;
        else start = live.front().off;
    }
    size_t in_use() const {
        size_t s = 0;
        for (auto &l : live) s += l.bytes;
        return s;
    }
};

struct Request {
    int64_t ticket = 0, q = 0;
    std::vector<float> x;
    // submitted, admitted to a kNN batch, kNN seen finished, layout launched, done
    double t_submit = 0, t_admit = 0, t_knn = 0, t_join = 0, t_done = 0;
    // layout state (device, in the arena)
    size_t arena_off = 0;
    int64_t E = 0;
    bool small = true;
    int32_t *offsets = nullptr, *indices = nullptr, *ranks = nullptr;
    double *per = nullptr, *npos = nullptr, *nper = nullptr, *nneg = nullptr;
    float *pos[2] = {nullptr, nullptr}, *m = nullptr, *v = nullptr, *proj = nullptr;
    unsigned long long *status = nullptr;
    int cur = 0;                 // buffer that holds the result
    hipEvent_t ev_final = nullptr;
    // result
    TGraphState *graph = nullptr;   // device: the request's graph state
    std::vector<float> out;
    umaphip_status code = UMAPHIP_OK;
    std::string error;
    bool done = false;         // guarded by the server mutex
};

// Device bytes of one request's layout state (graph of at most q*k edges).
size_t request_bytes(int64_t q, int64_t k) {
    const size_t E = static_cast<size_t>(q * k);
    return align256((q + 1) * 4) + align256(E * 4) + 4 * align256(E * 8) +
           4 * align256(q * 2 * 4) + align256(q * 4) + align256(q * 4) + 256 +
           align256(sizeof(TGraphState));
}

}  // namespace
}  // namespace umaphip

using namespace umaphip;

struct umaphip_server {
    // model
    int64_t n = 0, D = 0, d = 2, k = 15, epochs = 0;
    float a = 0, b = 0, gamma = 1, alpha0 = 0.25f, local_connectivity = 0, mix = 1;
    int32_t rate = 5, exclude = 0;
    uint32_t range = 0;
    float scale = 1;
    std::unique_ptr<KnnIndex> index;
    DeviceArray<float> tail, dirs;
    std::vector<float> tail_host, dirs_host;
    DeviceArray<int32_t> iota, cache_order;
    DeviceArray<float> cache_sorted, cache_table;
    std::vector<float> alpha, c1, c2;
    DeviceArray<float> sched;          // [3, epochs]: alpha, c1, c2
    // streams and staging
    static constexpr int kLayoutStreams = 3;
    hipStream_t sk = nullptr, sl[kLayoutStreams] = {};
    int64_t max_batch = 0;
    DeviceArray<float> q_dev, d_dist;
    DeviceArray<int32_t> d_idx;
    std::unique_ptr<Pinned> q_pin;
    hipEvent_t ev_knn = nullptr, ev_graph = nullptr;   // kNN outputs written / read
    TGraphParams gparams{};
    DeviceArray<int32_t> g_tail, g_cnt;               // graph workspace of one batch
    DeviceArray<float> g_w;
    // segment tables of in-flight layout launches (a ring; a slot is reused
    // once its launch has finished)
    static constexpr int kSlots = 8;
    std::unique_ptr<Pinned> seg_pin[kSlots], gseg_pin[kSlots];
    DeviceArray<TransformSegment> seg_dev[kSlots];
    DeviceArray<TGraphSegment> gseg_dev[kSlots];
    hipEvent_t ev_slot[kSlots] = {};
    int64_t launch_index = 0;
    DeviceArray<uint8_t> arena_mem;
    FifoArena arena;
    // queues
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::vector<std::shared_ptr<Request>> knn_batch;
    std::deque<std::shared_ptr<Request>> finishing;   // admission order
    std::unordered_map<int64_t, std::shared_ptr<Request>> tickets;
    int64_t next_ticket = 1;
    bool stop = false;
    std::string fatal;
    std::thread worker;
    // statistics
    int64_t st_requests = 0, st_rows = 0, st_launches = 0, st_knn_batches = 0, st_seg_sum = 0,
            st_peak_active = 0, st_rejected_wait = 0;
    double st_t_queue = 0, st_t_knn = 0, st_t_graph = 0, st_t_layout = 0;   // summed seconds

    ~umaphip_server() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
        for (auto st : sl)
            if (st) (void)hipStreamSynchronize(st);
        if (sk) (void)hipStreamSynchronize(sk);
        for (auto &e : ev_slot)
            if (e) (void)hipEventDestroy(e);
        if (ev_knn) (void)hipEventDestroy(ev_knn);
        if (ev_graph) (void)hipEventDestroy(ev_graph);
        for (auto &r : finishing)
            if (r->ev_final) (void)hipEventDestroy(r->ev_final);
        for (auto st : sl)
            if (st) (void)hipStreamDestroy(st);
        if (sk) (void)hipStreamDestroy(sk);
    }

    int device = 0;

    // Marks a request as failed; it retires through the finishing queue.
    void fail_request(const std::shared_ptr<Request> &r, umaphip_status code, const std::string &m) {
        r->code = code;
        r->error = m;
    }

    // ---- admission: a kNN batch of pending requests that fit the arena
    void start_knn_batch() {
        std::vector<std::shared_ptr<Request>> batch;
        int64_t rows = 0;
        {
            std::lock_guard<std::mutex> g(mu);
            while (!pending.empty()) {
                auto &r = pending.front();
                if (rows + r->q > max_batch) break;
                size_t off;
                if (!arena.alloc(request_bytes(r->q, k), off)) {
                    ++st_rejected_wait;
                    break;
                }
                r->arena_off = off;
                r->t_admit = now_seconds();
                rows += r->q;
                batch.push_back(r);
                pending.pop_front();
            }
        }
        if (batch.empty()) return;
        float *qp = static_cast<float *>(q_pin->ptr);
        int64_t row = 0;
        for (auto &r : batch) {
            std::memcpy(qp + row * D, r->x.data(), r->q * D * sizeof(float));
            row += r->q;
        }
        HIP_CHECK(hipMemcpyAsync(q_dev.ptr, qp, rows * D * sizeof(float), hipMemcpyHostToDevice, sk));
        // the previous batch's graph kernels have read the kNN outputs
        HIP_CHECK(hipStreamWaitEvent(sk, ev_graph, 0));
        index->query(q_dev.ptr, rows, static_cast<int>(k), false, 0, d_idx.ptr, d_dist.ptr, sk);
        HIP_CHECK(hipEventRecord(ev_knn, sk));
        knn_batch = std::move(batch);
        ++st_knn_batches;
    }

    // ---- device graphs and one layout launch for a finished kNN batch
    void join_knn_batch() {
        const int slot = static_cast<int>(launch_index % kSlots);
        hipStream_t ls = sl[launch_index % kLayoutStreams];
        HIP_CHECK(hipEventSynchronize(ev_slot[slot]));  // the slot's previous launch finished
        const double t_knn = now_seconds();
        const int32_t ns = static_cast<int32_t>(knn_batch.size());
        TGraphSegment *gsegs = static_cast<TGraphSegment *>(gseg_pin[slot]->ptr);
        TransformSegment *segs = static_cast<TransformSegment *>(seg_pin[slot]->ptr);
        int64_t row = 0, tiles = 0;
        for (int32_t i = 0; i < ns; ++i) {
            Request &r = *knn_batch[i];
            // carve the request's arena block
            uint8_t *p = arena.base + r.arena_off;
            auto take = [&](size_t bytes) {
                uint8_t *q = p;
                p += align256(bytes);
                return q;
            };
            r.small = r.q <= 1024;
            r.offsets = reinterpret_cast<int32_t *>(take((r.q + 1) * 4));
            r.indices = reinterpret_cast<int32_t *>(take(r.q * k * 4));
            r.per = reinterpret_cast<double *>(take(r.q * k * 8));
            r.npos = reinterpret_cast<double *>(take(r.q * k * 8));
            r.nper = reinterpret_cast<double *>(take(r.q * k * 8));
            r.nneg = reinterpret_cast<double *>(take(r.q * k * 8));
            r.pos[0] = reinterpret_cast<float *>(take(r.q * d * 4));
            r.pos[1] = reinterpret_cast<float *>(take(r.q * d * 4));
            r.m = reinterpret_cast<float *>(take(r.q * d * 4));
            r.v = reinterpret_cast<float *>(take(r.q * d * 4));
            r.ranks = reinterpret_cast<int32_t *>(take(r.q * 4));
            r.proj = reinterpret_cast<float *>(take(r.q * 4));
            r.status = reinterpret_cast<unsigned long long *>(take(8));
            r.graph = reinterpret_cast<TGraphState *>(take(sizeof(TGraphState)));
            r.cur = static_cast<int>(epochs & 1);
            TGraphSegment G{};
            G.row0 = row;
            G.rows = r.q;
            G.offsets = reinterpret_cast<uint32_t *>(r.offsets);
            G.indices = reinterpret_cast<uint32_t *>(r.indices);
            G.periods = r.per;
            G.negative_periods = r.nper;
            G.next_positive = r.npos;
            G.next_negative = r.nneg;
            G.init = r.pos[0];
            G.state = r.graph;
            gsegs[i] = G;
            row += r.q;
            SweepArgs A{};
            A.h = r.q; A.n = n;
            A.pos = r.pos[0]; A.pos_out = r.pos[1];
            A.tgt = tail.ptr;
            A.offsets = r.offsets; A.indices = r.indices;
            A.periods = r.per; A.next_positive = r.npos;
            A.negative_periods = r.nper; A.next_negative = r.nneg;
            A.m = r.m; A.v = r.v;
            A.scale = scale;
            A.a = a; A.b = b; A.gamma = gamma; A.range = range;
            A.multiplier = 1.0f;
            A.shared = 0; A.negatives = rate != 0; A.exclude = exclude; A.fused = 1;
            segs[i].args = A;
            segs[i].tile0 = tiles;
            segs[i].ranks = r.ranks;
            segs[i].status = r.status;
            tiles += (r.q + transform_tile_rows() - 1) / transform_tile_rows();
        }
        HIP_CHECK(hipMemcpyAsync(gseg_dev[slot].ptr, gsegs, sizeof(TGraphSegment) * ns,
                                 hipMemcpyHostToDevice, ls));
        HIP_CHECK(hipStreamWaitEvent(ls, ev_knn, 0));
        launch_transform_graphs(gseg_dev[slot].ptr, ns, row, d_idx.ptr, d_dist.ptr, k, n, gparams,
                                tail.ptr, d, static_cast<int32_t>(std::max<int64_t>(epochs, 1)),
                                static_cast<float>(rate), g_tail.ptr, g_w.ptr, g_cnt.ptr, ls);
        HIP_CHECK(hipEventRecord(ev_graph, ls));
        for (int32_t i = 0; i < ns; ++i) {
            Request &r = *knn_batch[i];
            HIP_CHECK(hipMemsetAsync(r.m, 0, r.q * d * 4, ls));
            HIP_CHECK(hipMemsetAsync(r.v, 0, r.q * d * 4, ls));
            launch_status_min_u64(r.status, ~0ull, ls);
        }
        if (epochs > 0) {
            HIP_CHECK(hipMemcpyAsync(seg_dev[slot].ptr, segs, sizeof(TransformSegment) * ns,
                                     hipMemcpyHostToDevice, ls));
            TransformTables T{};
            T.alpha = sched.ptr;
            T.c1 = sched.ptr + epochs;
            T.c2 = sched.ptr + 2 * epochs;
            T.iota = iota.ptr;
            T.tail = tail.ptr;
            T.order = cache_order.ptr;
            T.sorted = cache_sorted.ptr;
            T.table = cache_table.ptr;
            T.dirs = dirs.ptr;
            T.n = n;
            T.epochs = static_cast<int32_t>(epochs);
            launch_transform_layouts(seg_dev[slot].ptr, ns, tiles, T, ls);
        }
        HIP_CHECK(hipEventRecord(ev_slot[slot], ls));
        ++launch_index;
        const double t_join = now_seconds();
        for (auto &r : knn_batch) {
            r->t_knn = t_knn;
            r->t_join = t_join;
            HIP_CHECK(hipEventCreateWithFlags(&r->ev_final, hipEventDisableTiming));
            HIP_CHECK(hipEventRecord(r->ev_final, ls));
        }
        std::lock_guard<std::mutex> g(mu);
        for (auto &r : knn_batch) finishing.push_back(r);
        ++st_launches;
        st_seg_sum += ns;
        st_peak_active = std::max<int64_t>(st_peak_active, static_cast<int64_t>(finishing.size()));
        knn_batch.clear();
    }

    // ---- retire every finished request (launches on different streams may
    // finish out of admission order; the arena frees blocks by offset)
    bool retire() {
        std::vector<std::shared_ptr<Request>> ready;
        {
            std::lock_guard<std::mutex> g(mu);
            for (auto it = finishing.begin(); it != finishing.end();) {
                if (hipEventQuery((*it)->ev_final) == hipSuccess) {
                    ready.push_back(*it);
                    it = finishing.erase(it);
                } else {
                    ++it;
                }
            }
        }
        if (ready.empty()) return false;
        for (auto &r : ready) {
            (void)hipEventDestroy(r->ev_final);
            r->ev_final = nullptr;
            if (r->code == UMAPHIP_OK) {
                TGraphState gs{};
                HIP_CHECK(hipMemcpy(&gs, r->graph, sizeof(gs), hipMemcpyDeviceToHost));
                if (gs.errors) {
                    fail_request(r, transform_graph_status(gs.errors), transform_graph_error(gs.errors));
                } else {
                    r->out.resize(r->q * d);
                    unsigned long long status = ~0ull;
                    HIP_CHECK(hipMemcpy(r->out.data(), r->pos[r->cur], r->q * d * sizeof(float),
                                        hipMemcpyDeviceToHost));
                    HIP_CHECK(hipMemcpy(&status, r->status, sizeof(status), hipMemcpyDeviceToHost));
                    bool finite = status == ~0ull;
                    for (float x : r->out) finite = finite && std::isfinite(x);
                    if (!finite) fail_request(r, UMAPHIP_NUMERICAL, "non-finite transform coordinates");
                }
            }
        }
        {
            std::lock_guard<std::mutex> g(mu);
            const double t = now_seconds();
            for (auto &r : ready) {
                r->t_done = t;
                st_t_queue += r->t_admit - r->t_submit;
                st_t_knn += r->t_knn - r->t_admit;
                st_t_graph += r->t_join - r->t_knn;
                st_t_layout += r->t_done - r->t_join;
                r->done = true;
                arena.release(r->arena_off);
                ++st_requests;
                st_rows += r->q;
            }
        }
        cv_done.notify_all();
        return true;
    }

    void loop() {
        (void)hipSetDevice(device);
        try {
            for (;;) {
                bool busy = false;
                {
                    std::unique_lock<std::mutex> g(mu);
                    if (stop) break;
                    if (pending.empty() && knn_batch.empty() && finishing.empty())
                        cv_work.wait_for(g, std::chrono::milliseconds(5));
                    if (stop) break;
                }
                if (knn_batch.empty()) start_knn_batch();
                if (!knn_batch.empty() && hipEventQuery(ev_knn) == hipSuccess) {
                    join_knn_batch();
                    busy = true;
                }
                busy = retire() || busy;
                if (!busy && (!finishing.empty() || !knn_batch.empty()))
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
        } catch (const Fail &) {
            std::lock_guard<std::mutex> g(mu);
            fatal = umaphip_last_error();
            auto fail_all = [&](auto &list) {
                for (auto &r : list) {
                    r->code = UMAPHIP_BACKEND;
                    r->error = fatal;
                    r->done = true;
                }
            };
            fail_all(knn_batch);
            fail_all(pending);
            fail_all(finishing);
            stop = true;
            cv_done.notify_all();
        }
    }
};

extern "C" {

umaphip_status umaphip_server_create(const float *reference, int64_t n, int64_t features,
                                     const float *embedding, int64_t d,
                                     const umaphip_server_params *params, umaphip_server **out) {
    return guarded([&] {
        UMAPHIP_REQUIRE(out && reference && embedding && params, UMAPHIP_INVALID_INPUT,
                        "server arguments are null");
        *out = nullptr;
        UMAPHIP_REQUIRE(n > 0 && features > 0 && n <= INT32_MAX - 1, UMAPHIP_INVALID_INPUT,
                        "server needs a non-empty reference set");
        UMAPHIP_REQUIRE(d == 2, UMAPHIP_INVALID_CONFIG, "the transform server supports 2-D embeddings");
        const umaphip_server_params &p = *params;
        UMAPHIP_REQUIRE(p.n_neighbors >= 1 && p.epochs >= 0 && p.epochs < 65535 &&
                            p.negative_sample_rate >= 0 && p.negative_selection_range > 0 &&
                            p.max_batch_rows >= 1,
                        UMAPHIP_INVALID_CONFIG, "invalid server parameters");
        std::unique_ptr<umaphip_server> S(new umaphip_server());
        S->n = n; S->D = features; S->d = d;
        S->k = std::min<int64_t>(p.n_neighbors, n);
        UMAPHIP_REQUIRE(S->k <= 64, UMAPHIP_INVALID_CONFIG, "n_neighbors above 64 is not supported");
        S->epochs = p.epochs;
        S->a = p.a; S->b = p.b; S->gamma = p.repulsion;
        S->alpha0 = p.learning_rate / 4.0f;
        S->local_connectivity = std::max(p.local_connectivity - 1.0f, 0.0f);
        S->mix = p.set_op_mix_ratio;
        S->rate = p.negative_sample_rate;
        S->exclude = p.exclude_graph_neighbors;
        const uint64_t Q = static_cast<uint64_t>(p.negative_selection_range);
        UMAPHIP_REQUIRE(Q < static_cast<uint64_t>(n), UMAPHIP_INVALID_CONFIG,
                        "the server caches projection windows: negative_selection_range must be "
                        "below the reference rows");
        S->range = static_cast<uint32_t>(effective_range(Q, static_cast<uint64_t>(n)));
        S->scale = resolved_scale(Q, static_cast<uint64_t>(n), p.has_scale != 0, p.scale,
                                  p.adaptation_samples > 0);
        std::vector<float> mu;
        schedules(static_cast<int>(S->epochs), S->alpha0, false, false, S->alpha, mu);
        bias_corrections(static_cast<int>(S->epochs), S->c1, S->c2);
        S->max_batch = p.max_batch_rows;
        for (auto &st : S->sl) HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIP_CHECK(hipStreamCreateWithFlags(&S->sk, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreateWithFlags(&S->ev_knn, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&S->ev_graph, hipEventDisableTiming));
        HIP_CHECK(hipEventRecord(S->ev_graph, S->sl[0]));
        for (int w = 0; w < umaphip_server::kSlots; ++w) {
            HIP_CHECK(hipEventCreateWithFlags(&S->ev_slot[w], hipEventDisableTiming));
            HIP_CHECK(hipEventRecord(S->ev_slot[w], S->sl[0]));
            // a batch holds at most max_batch requests (each has a row)
            S->seg_pin[w].reset(new Pinned(sizeof(TransformSegment) * S->max_batch));
            S->seg_dev[w].alloc(S->max_batch);
            S->gseg_pin[w].reset(new Pinned(sizeof(TGraphSegment) * S->max_batch));
            S->gseg_dev[w].alloc(S->max_batch);
        }
        if (S->epochs > 0) {
            std::vector<float> sched(S->alpha);
            sched.insert(sched.end(), S->c1.begin(), S->c1.end());
            sched.insert(sched.end(), S->c2.begin(), S->c2.end());
            S->sched.alloc(sched.size());
            HIP_CHECK(hipMemcpy(S->sched.ptr, sched.data(), sched.size() * sizeof(float),
                                hipMemcpyHostToDevice));
        }
        // resident model
        S->index.reset(new KnnIndex(reference, nullptr, n, features, S->max_batch, S->sk));
        S->tail_host.assign(embedding, embedding + n * d);
        S->tail.alloc(n * d);
        HIP_CHECK(hipMemcpy(S->tail.ptr, embedding, n * d * sizeof(float), hipMemcpyHostToDevice));
        S->iota.alloc(n);
        launch_iota(S->iota.ptr, n, S->sk);
        // per-direction projection tables of the fixed tail (directions of
        // StdRng(transform_seed), one per epoch)
        const int64_t nd = std::max<int64_t>(S->epochs, 1);
        S->dirs_host.resize(nd * d);
        StdRng rng(p.transform_seed);
        for (int64_t t = 0; t < nd; ++t) projection_direction(rng, d, S->dirs_host.data() + t * d);
        S->dirs.alloc(nd * d);
        HIP_CHECK(hipMemcpy(S->dirs.ptr, S->dirs_host.data(), nd * d * sizeof(float),
                            hipMemcpyHostToDevice));
        S->cache_order.alloc(nd * n);
        S->cache_sorted.alloc(nd * n);
        S->cache_table.alloc(nd * n * 2);
        {
            DeviceArray<uint32_t> keys(n), keys_sorted(n);
            DeviceArray<uint8_t> ws(sort_pairs_workspace(n));
            DeviceArray<unsigned long long> status(1);
            launch_status_min_u64(status.ptr, ~0ull, S->sk);
            for (int64_t t = 0; t < nd; ++t) {
                launch_project_keys(S->tail.ptr, n, d, S->dirs.ptr + t * d, keys.ptr, nullptr,
                                    status.ptr, static_cast<unsigned long long>(t) << 32, S->sk);
                launch_sort_pairs(ws.ptr, ws.count, keys.ptr, keys_sorted.ptr, S->iota.ptr,
                                  S->cache_order.ptr + t * n, n, S->sk);
                launch_values_from_keys(keys_sorted.ptr, n, S->cache_sorted.ptr + t * n, S->sk);
                launch_gather_table(S->cache_order.ptr + t * n, n, S->tail.ptr,
                                    S->cache_table.ptr + t * n * 2, S->sk);
            }
            unsigned long long st = 0;
            HIP_CHECK(hipMemcpyAsync(&st, status.ptr, sizeof(st), hipMemcpyDeviceToHost, S->sk));
            HIP_CHECK(hipStreamSynchronize(S->sk));
            UMAPHIP_REQUIRE(st == ~0ull, UMAPHIP_NUMERICAL, "non-finite fitted embedding at row %llu",
                            st & 0xFFFFFFFFull);
        }
        // kNN staging
        S->q_dev.alloc(S->max_batch * features);
        S->d_idx.alloc(S->max_batch * S->k);
        S->d_dist.alloc(S->max_batch * S->k);
        S->q_pin.reset(new Pinned(S->max_batch * features * sizeof(float)));
        // device graph workspace of one batch and the graph settings
        // (transform local connectivity max(lc - 1, 0), bandwidth 1)
        S->g_tail.alloc(S->max_batch * S->k);
        S->g_w.alloc(S->max_batch * S->k);
        S->g_cnt.alloc(S->max_batch);
        S->gparams = transform_graph_params(S->k, S->local_connectivity);
        // admission arena: explicit size, or a quarter of the free HBM up to 1 GiB
        size_t free_b = 0, total_b = 0;
        HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
        size_t cap = p.memory_bytes > 0 ? static_cast<size_t>(p.memory_bytes)
                                        : std::min<size_t>(free_b / 4, size_t(1) << 30);
        UMAPHIP_REQUIRE(cap >= request_bytes(1, S->k) && cap <= free_b, UMAPHIP_MEMORY,
                        "server memory budget %zu exceeds free device memory %zu", cap, free_b);
        S->arena_mem.alloc(cap);
        S->arena.base = S->arena_mem.ptr;
        S->arena.cap = cap;
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipGetDevice(&S->device));
        umaphip_server *raw = S.get();
        S->worker = std::thread([raw] { raw->loop(); });
        *out = S.release();
    });
}

umaphip_status umaphip_server_submit(umaphip_server *server, const float *queries, int64_t q,
                                     int64_t *ticket) {
    return guarded([&] {
        UMAPHIP_REQUIRE(server && ticket && (q == 0 || queries), UMAPHIP_INVALID_INPUT,
                        "server submit arguments are null");
        UMAPHIP_REQUIRE(q >= 1 && q <= server->max_batch, UMAPHIP_CAPACITY,
                        "request rows must be in 1..=%lld", (long long)server->max_batch);
        UMAPHIP_REQUIRE(request_bytes(q, server->k) <= server->arena.cap, UMAPHIP_MEMORY,
                        "request exceeds the server memory budget");
        auto r = std::make_shared<Request>();
        r->q = q;
        r->x.assign(queries, queries + q * server->D);
        for (float v : r->x)
            UMAPHIP_REQUIRE(std::isfinite(v), UMAPHIP_INVALID_INPUT, "query rows must be finite");
        r->t_submit = now_seconds();
        {
            std::lock_guard<std::mutex> g(server->mu);
            UMAPHIP_REQUIRE(server->fatal.empty(), UMAPHIP_BACKEND, "server stopped: %s",
                            server->fatal.c_str());
            r->ticket = server->next_ticket++;
            server->tickets[r->ticket] = r;
            server->pending.push_back(r);
        }
        server->cv_work.notify_all();
        *ticket = r->ticket;
    });
}

umaphip_status umaphip_server_wait(umaphip_server *server, int64_t ticket, double timeout_s,
                                   float *out, double *latency_ms, int32_t *done) {
    return guarded([&] {
        UMAPHIP_REQUIRE(server && done, UMAPHIP_INVALID_INPUT, "server wait arguments are null");
        *done = 0;
        std::shared_ptr<Request> r;
        std::unique_lock<std::mutex> g(server->mu);
        auto it = server->tickets.find(ticket);
        UMAPHIP_REQUIRE(it != server->tickets.end(), UMAPHIP_INVALID_INPUT, "unknown ticket %lld",
                        (long long)ticket);
        r = it->second;
        const auto deadline = std::chrono::steady_clock::now() +
                              std::chrono::duration<double>(std::max(timeout_s, 0.0));
        while (!r->done) {
            if (timeout_s <= 0.0 ||
                server->cv_done.wait_until(g, deadline) == std::cv_status::timeout) {
                if (!r->done) break;
            }
        }
        *done = r->done ? 1 : 0;
        if (!r->done) return;
        server->tickets.erase(ticket);
        g.unlock();
        if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
        UMAPHIP_REQUIRE(r->code == UMAPHIP_OK, r->code, "%s", r->error.c_str());
        if (out) std::memcpy(out, r->out.data(), r->out.size() * sizeof(float));
    });
}

umaphip_status umaphip_server_stats(umaphip_server *server, umaphip_server_counters *stats) {
    return guarded([&] {
        UMAPHIP_REQUIRE(server && stats, UMAPHIP_INVALID_INPUT, "server stats arguments are null");
        std::lock_guard<std::mutex> g(server->mu);
        stats->requests = server->st_requests;
        stats->rows = server->st_rows;
        stats->steps = server->st_launches;
        stats->knn_batches = server->st_knn_batches;
        stats->mean_active = server->st_launches ? static_cast<double>(server->st_seg_sum) /
                                                       static_cast<double>(server->st_launches)
                                                 : 0.0;
        stats->peak_active = server->st_peak_active;
        stats->admission_waits = server->st_rejected_wait;
        stats->memory_budget = static_cast<int64_t>(server->arena.cap);
        stats->memory_in_use = static_cast<int64_t>(server->arena.in_use());
        const double nreq = server->st_requests ? static_cast<double>(server->st_requests) : 1.0;
        stats->mean_queue_ms = server->st_t_queue / nreq * 1e3;
        stats->mean_knn_ms = server->st_t_knn / nreq * 1e3;
        stats->mean_graph_ms = server->st_t_graph / nreq * 1e3;
        stats->mean_layout_ms = server->st_t_layout / nreq * 1e3;
    });
}

void umaphip_server_destroy(umaphip_server *server) { delete server; }

}  // extern "C"
