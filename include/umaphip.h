/*
 * umaphip -- MI355X-native modern UMAP layout path behind a C ABI.
 *
 * Every entry point takes plain pointers and sizes. Device pointers are
 * HIP device allocations; `stream` is a hipStream_t passed as void* (NULL =
 * the null stream). Calls enqueue work in stream order and do not
 * synchronise unless stated. No entry point allocates device memory except
 * the plan/server constructors and the host-array drivers, which say so.
 *
 * Errors: every function returns a umaphip_status; umaphip_last_error()
 * returns the calling thread's last message. The taxonomy mirrors cuscan's
 * UmapError (crates/umap-core/src/lib.rs:26-58) and its Python mapping
 * (crates/cuscan-python/src/lib.rs:373-384): INVALID_INPUT / INVALID_CONFIG
 * -> ValueError, MEMORY -> MemoryError, the rest -> RuntimeError.
 */
#ifndef UMAPHIP_H
#define UMAPHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    UMAPHIP_OK = 0,
    UMAPHIP_INVALID_INPUT = 1,
    UMAPHIP_INVALID_CONFIG = 2,
    UMAPHIP_MEMORY = 3,
    UMAPHIP_CAPACITY = 4,
    UMAPHIP_NUMERICAL = 5,
    UMAPHIP_BACKEND = 6,
    UMAPHIP_OVERFLOW = 7
} umaphip_status;

/* Thread-local message of the last failing call on this thread. */
const char *umaphip_last_error(void);
/* Library ABI version (major * 10000 + minor * 100 + patch). */
int32_t umaphip_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
int32_t umaphip_device_count(void);

/* ------------------------------------------------------------------------
 * One modern epoch. Replaces StandardUmapOps::modern_epoch
 * (crates/umap/src/backend.rs:144-232; CubeCL body
 * crates/umap/src/kernels/standard_umap/modern_layout.rs:925-1026).
 *
 * Arguments mean what the reference's do. Device pointers; 1-D arrays are
 * contiguous; coordinate arrays use (row, column) strides in elements.
 * head, next_positive, next_negative, updates, m and v are overwritten in
 * place and must not alias each other or the read-only inputs. `shared`
 * selects fit (targets read from head); otherwise fixed_tail is read-only.
 * Adam: the gradient is accumulated into updates (which therefore should
 * hold zeros on entry) and one lazy Adam step consumes and clears it.
 * Momentum: updates is the velocity; source_order is split into blocks of
 * H (H <= 1024) or H/2 rows, each block's gradient (times alpha) is added
 * to the velocity and applied before the next block; the velocity is then
 * decayed by mu. Uses no scratch memory.
 * ------------------------------------------------------------------------ */
typedef struct {
    int64_t h, n, d;
    float *head;
    int64_t head_s0, head_s1;
    const float *fixed_tail;
    int64_t tail_s0, tail_s1;
    const int32_t *offsets;        /* h + 1 */
    const int32_t *indices;        /* E, sorted within each row */
    const int32_t *source_order;   /* h, permutation of rows */
    const int32_t *negative_order; /* n, permutation of reference rows */
    const int32_t *source_ranks;   /* h, read only when negative_range < n and h > 1024 */
    const double *periods;
    double *next_positive;
    const double *negative_periods;
    double *next_negative;
    float *updates;
    int64_t upd_s0, upd_s1;
    float *m;
    int64_t m_s0, m_s1;
    float *v;
    int64_t v_s0, v_s1;
    uint32_t epoch;
    float a, b, gamma, scale;
    uint32_t negative_range;      /* effective window R in 1..=n */
    float alpha, mu, correction1, correction2;
    int32_t momentum, shared, negatives, exclude_graph_neighbors;
    int64_t n_edges;              /* E = offsets[h] (host-known; checked by debug entry) */
} umaphip_epoch_args;

umaphip_status umaphip_modern_epoch(const umaphip_epoch_args *args, void *stream);

/* Test hook (no reference counterpart): writes, for every edge of the
 * epoch's graph, the negative candidates the gradient would visit at this
 * epoch (count_out[e] = number visited, -1 when the edge does not fire;
 * cand_out[e*cap + k] = candidate row, kept_out[e*cap + k] = 1 when not
 * excluded). Runs the same device selection code as the gradient kernels.
 * Reads but does not modify the clocks. Fails with CAPACITY when an edge
 * visits more than `cap` candidates. Synchronises the stream. */
umaphip_status umaphip_debug_negatives(const umaphip_epoch_args *args, int32_t cap,
                                       int32_t *count_out, int32_t *cand_out,
                                       uint8_t *kept_out, void *stream);

/* ------------------------------------------------------------------------
 * FP64 objective (no reference counterpart; the reference computes no
 * objective). Fuzzy-set cross entropy of the embedding over the graph's
 * positive edges:
 *   sum_e -w_e ln(max(phi_e, 1e-4)) - (1 - w_e) ln(max(1 - phi_e, 1e-4)),
 *   phi_e = 1 / (1 + a q_e^b), q_e = |head_i - tail_j|^2,
 * with every term and sum in FP64 and a deterministic reduction order.
 * tail = NULL for a fit (targets are head rows). Device pointers; *out is
 * a device double.
 * ------------------------------------------------------------------------ */
umaphip_status umaphip_cross_entropy_workspace_bytes(int64_t h, size_t *bytes);
umaphip_status umaphip_cross_entropy(const float *head, int64_t h, const float *tail, int64_t n,
                                     int64_t d, const int32_t *offsets, const int32_t *indices,
                                     const float *weights, float a, float b, void *workspace,
                                     size_t workspace_bytes, double *out, void *stream);

/* ------------------------------------------------------------------------
 * Projection windows. Replace modern_project / modern_projection_ranks
 * (kernels/standard_umap/modern_layout.rs:669-752) and the stable f32
 * argsort (crates/cuscan-common/src/radix.rs:659, key :993-1003).
 * ------------------------------------------------------------------------ */
/* out[i] = sequential f32 sum_c x[i*s0 + c*s1] * direction[c]. */
umaphip_status umaphip_project(const float *x, int64_t rows, int64_t d, int64_t s0, int64_t s1,
                               const float *direction, float *out, void *stream);
/* Workspace bytes needed by umaphip_stable_argsort_f32 for n keys. */
umaphip_status umaphip_argsort_workspace_bytes(int64_t n, size_t *bytes);
/* Stable ascending argsort of finite floats (-0 == +0). order[j] = row of
 * rank j, sorted[j] = values[order[j]] (+0 for a signed zero). status_out
 * (device, one int64) receives -1 or the lowest non-finite row. */
umaphip_status umaphip_stable_argsort_f32(const float *values, int64_t n, int32_t *order,
                                          float *sorted, int64_t *status_out, void *workspace,
                                          size_t workspace_bytes, void *stream);
/* umaphip_stable_argsort_f32 without a workspace argument and without any
 * allocation. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation). Its workspace is a
 * static device array of about 20.5 MB per device, held until the process
 * ends, so n is at most 1,048,576 keys (UMAPHIP_CAPACITY beyond that).
 * Results equal umaphip_stable_argsort_f32's. Every call on a device
 * shares that device's array: calls that overlap on one device (on
 * different streams or host threads) race and must be serialised. */
umaphip_status umaphip_stable_argsort_f32_noalloc(const float *values, int64_t n, int32_t *order,
                                                  float *sorted, int64_t *status_out,
                                                  void *stream);
/* shared: ranks[order[j]] = j (n entries). Otherwise ranks[i] =
 * min(lower_bound(sorted, query[i]), n - 1) for h queries. */
umaphip_status umaphip_projection_ranks(const int32_t *order, const float *sorted, int64_t n,
                                        const float *query, int64_t h, int32_t *ranks,
                                        int32_t shared, void *stream);

/* Adaptation resultants. Replaces modern_sample_resultants
 * (kernels/standard_umap/modern_layout.rs:754-823): 2*S float64 norms,
 * global at [0,S), local at [S,2S). embedding is [n, d] contiguous. */
umaphip_status umaphip_sample_resultants(const float *embedding, int64_t n, int64_t d,
                                         const int32_t *order, const int32_t *offsets,
                                         const int32_t *indices, const int32_t *sample_ranks,
                                         const int32_t *local_offsets,
                                         const int32_t *global_ranks, int64_t samples,
                                         uint32_t raw_range, float a, float b, float gamma,
                                         int32_t exclude, double *resultants, void *stream);

/* ------------------------------------------------------------------------
 * Layout driver. Replaces modern_runtime::optimize_compact /
 * optimize_compact_inner (crates/umap/src/standard_umap/modern_runtime.rs:
 * 1011-1378): orders, projection refreshes, adaptation and the epoch loop,
 * device resident with no per-epoch host synchronisation.
 * ------------------------------------------------------------------------ */
typedef struct {
    int32_t optimizer;              /* 0 = adam, 1 = momentum */
    int32_t epochs;
    float a, b, repulsion, initial_alpha;
    int64_t negative_sample_rate;
    int64_t negative_selection_range;
    int32_t has_scale;
    float scale;
    int64_t adaptation_samples;
    int32_t exclude_graph_neighbors;
    int32_t good_initialization;
    uint64_t seed;
} umaphip_layout_params;

typedef struct {
    int64_t epochs_run;
    int64_t adaptations;
    int64_t refreshes;
    float final_scale;
    double device_ms;               /* epoch loop, device events */
    double sweep_ms;                /* sum over edge-sweep launches (profiling on) */
    int64_t sweep_launches;
    double refresh_ms;              /* sum over projection refreshes (profiling on) */
    int64_t refresh_launches;
    int64_t edge_balanced;          /* 1 when epochs ran the edge-balanced unit sweep */
    int64_t rank_ordered_epochs;    /* epochs of the last run whose unit sweep ran the rows
                                       in projection order */
    int64_t compact_records;        /* 1 when the rank-ordered sweep kept 24-byte edge records
                                       (every period rebuilt exactly from an f32 sample count) */
    int64_t clock_schedule;         /* epochs per clock-schedule pass of the edge-balanced sweep,
                                       0 when the sweep advances the clocks itself */
} umaphip_layout_stats;

/* Host arrays in, host arrays out (allocates and frees device memory).
 * head [h,d] in/out; tail [n,d] read for transform (ignored when shared);
 * CSR offsets [h+1], indices [E]; clocks [E] in/out. */
umaphip_status umaphip_optimize_compact(int64_t h, int64_t n, int64_t d, float *head,
                                        const float *tail, const uint32_t *offsets,
                                        const uint32_t *indices, int64_t n_edges,
                                        const double *periods, double *next_positive,
                                        const double *negative_periods, double *next_negative,
                                        const umaphip_layout_params *params, int32_t shared,
                                        umaphip_layout_stats *stats);

/* Resident plan: uploads once, then runs from device-resident initial
 * state any number of times (bench / serving). */
typedef struct umaphip_plan umaphip_plan;
umaphip_status umaphip_plan_create(int64_t h, int64_t n, int64_t d, const float *head,
                                   const float *tail, const uint32_t *offsets,
                                   const uint32_t *indices, int64_t n_edges,
                                   const double *periods, const double *next_positive,
                                   const double *negative_periods, const double *next_negative,
                                   const umaphip_layout_params *params, int32_t shared,
                                   umaphip_plan **out);
/* Objective trace (no reference counterpart): with every > 0, each run
 * evaluates the FP64 cross entropy (umaphip_cross_entropy) of the embedding
 * after every `every`-th epoch and after the last one, on the plan's
 * stream. weights [n_edges] are the CSR edge weights; every = 0 disables. */
umaphip_status umaphip_plan_set_objective(umaphip_plan *plan, const float *weights,
                                          int32_t every);
/* Synchronises; copies up to `capacity` values of the last run's trace. */
umaphip_status umaphip_plan_objective(umaphip_plan *plan, double *values, int64_t capacity,
                                      int64_t *count);
/* Resets the state to the uploaded initial state and enqueues every epoch
 * on the plan's stream. Does not synchronise. */
umaphip_status umaphip_plan_run(umaphip_plan *plan);
/* Waits for the plan's stream; copies the embedding (h*d floats) and, when
 * non-NULL, the clocks (E doubles each) to host. */
umaphip_status umaphip_plan_download(umaphip_plan *plan, float *head, double *next_positive,
                                     double *next_negative, umaphip_layout_stats *stats);
/* Records device events around every fused edge sweep and projection
 * refresh of later runs (per-kernel durations in the stats). */
umaphip_status umaphip_plan_set_profiling(umaphip_plan *plan, int32_t enabled);
/* Edge-record format of a plan on the edge-balanced sweep (no reference
 * counterpart; the reference keeps its clocks as four f64 arrays):
 * compact = 1 selects 24-byte records {next_positive, next_negative,
 * column | f32 sample count} (rank-ordered plans whose every period equals
 * epochs / f64(f32(epochs / period))), 0 selects 32-byte records that keep
 * the f64 period. Results are identical; plans default to 24 bytes when
 * allowed. */
umaphip_status umaphip_plan_set_record_format(umaphip_plan *plan, int32_t compact);
/* Clock schedule of a plan on the edge-balanced sweep (no reference counterpart; the
 * reference advances both clocks inside its gradient kernel,
 * kernels/standard_umap/modern_layout.rs:166-253). The clocks do not depend
 * on the positions, so a pass at the start of every span epochs runs them
 * ahead and stores each edge's term count per epoch (one byte per edge and
 * epoch); the sweep then reads that byte and the CSR column instead of the
 * edge record. span < 0 picks the span automatically (all epochs in one
 * pass while the bytes fit 4 GiB; the default), 0 turns the schedule off,
 * span > 0 caps the epochs per pass. Results are identical in every form.
 * The automatic schedule stays off when its memory cannot be allocated.
 * Needs a plan on the edge-balanced sweep (Adam, d = 2, more than 1024
 * rows) with negative_sample_rate <= 64 (any span > 0 on another plan is an
 * error; span < 0 leaves it off). */
umaphip_status umaphip_plan_set_clock_schedule(umaphip_plan *plan, int32_t span);
/* Epoch trace, a test hook with no reference counterpart: it stands where
 * the reference's runtime observers watch one epoch's state
 * (EPOCH_OBSERVER / FORCE_OBSERVER, crates/umap/src/standard_umap/
 * modern_runtime.rs:56-59, 1552-1939). For plans on the edge-balanced sweep
 * only. With epoch >= 0, later runs record, at that epoch, the state the
 * sweep starts from and what it produced; epoch < 0 turns it off.
 * negative_cap bounds the negatives recorded per edge. */
umaphip_status umaphip_plan_set_trace(umaphip_plan *plan, int32_t epoch, int32_t negative_cap);
/* Synchronises and copies the last run's trace (NULL outputs are skipped):
 * head_before [h*2] positions at the start of the epoch; order [n] the
 * negative order in force; ranks [h] the window ranks (identity when
 * negatives are global); scale [1] the repulsion scale in force;
 * gradient [h*2] each row's assembled gradient, the value the lazy Adam
 * step consumes; negative_count [E] the negatives drawn by each edge, -1
 * when it did not fire; negative_candidates / negative_kept [E*cap] the
 * candidate id of each draw and 1 unless exclusion dropped it. info [5]:
 * rank-ordered sweep, 24-byte records, unit row-table capacity, units on
 * the term-parallel path, units on the per-edge path. UMAPHIP_CAPACITY
 * when an edge drew more than negative_cap. */
umaphip_status umaphip_plan_trace(umaphip_plan *plan, float *head_before, int32_t *order,
                                  int32_t *ranks, float *scale, float *gradient,
                                  int32_t *negative_count, int32_t *negative_candidates,
                                  uint8_t *negative_kept, int32_t *info);
/* The plan's stream (hipStream_t as void*). */
void *umaphip_plan_stream(umaphip_plan *plan);
void umaphip_plan_destroy(umaphip_plan *plan);

/* Orders the driver would use (test hook for the host RNG): writes the
 * source and negative orders in force at `epoch` for a layout with the
 * given shapes/params, as the reference's ModernOrders would
 * (modern_orders.rs:359-680). */
umaphip_status umaphip_orders_at_epoch(int64_t h, int64_t n, const umaphip_layout_params *params,
                                       int32_t shared, int32_t epoch, int32_t *source_out,
                                       int32_t *negative_out);

/* ------------------------------------------------------------------------
 * Exact kNN. Replaces cuscan_common::neighbors::exact_neighbors
 * (crates/cuscan-common/src/neighbors.rs:148; kernels/neighbors.rs:270-560)
 * as reached from crates/umap/src/standard_umap/neighbors.rs:17-77. A BF16
 * MFMA distance prefilter with FP32 accumulation keeps 64 candidates per
 * query; an exact FP32 rerank then reproduces the reference's sequential
 * squared-distance sum, lower-index tie-break, square root and sentinels
 * (id = n, distance = sqrt(f32::MAX)); a certificate per query, with an
 * exact FP32 rescan where it fails, makes the result the reference's for
 * any data. k <= 64, features <= 1024.
 * ------------------------------------------------------------------------ */
typedef struct umaphip_knn_index umaphip_knn_index;
/* Copies n x d f32 references (host or device pointer) into a resident
 * index sized for batches of up to max_queries rows. Allocates. */
umaphip_status umaphip_knn_index_create(const float *reference, int32_t reference_on_device,
                                        int64_t n, int64_t d, int64_t max_queries,
                                        umaphip_knn_index **out);
/* queries [q, d] device f32; indices/distances [q, k] device outputs.
 * exclude_self skips reference row query_offset + i for query row i. */
umaphip_status umaphip_knn_query(umaphip_knn_index *index, const float *queries, int64_t q,
                                 int32_t k, int32_t exclude_self, int64_t query_offset,
                                 int32_t *indices, float *distances, void *stream);
/* Diagnostics (no reference counterpart): the BF16 prefilter's merged
 * candidate lists before the exact rerank, candidates [q, 64] best first by
 * approximate distance and counts [q] (device pointers); used to report the
 * prefilter's recall. */
umaphip_status umaphip_knn_candidates(umaphip_knn_index *index, const float *queries, int64_t q,
                                      int32_t *candidates, int32_t *counts, void *stream);
/* Exactness certificate (no reference counterpart; it keeps the prefilter
 * from changing the reference's results). Each query's top k is certified
 * against a bound on the BF16 rounding (the index subtracts a per-feature
 * shift near the references' mean before rounding) and the f32
 * accumulation of the approximate distances; a query the bound does not
 * certify is rescanned exactly in FP32 over every reference. every_query =
 * 1 rescans every query (tests), 0 restores the certificate. */
umaphip_status umaphip_knn_set_rescan(umaphip_knn_index *index, int32_t every_query);
/* Synchronises stream; rescanned (may be NULL) = queries of the last
 * umaphip_knn_query that were rescanned exactly; rounding_ratio (may be
 * NULL) = the largest BF16 rounding of a reference row relative to its norm
 * (0 when the shifted references are exact in BF16). */
umaphip_status umaphip_knn_certificate_stats(umaphip_knn_index *index, void *stream,
                                             int64_t *rescanned, float *rounding_ratio);
void umaphip_knn_index_destroy(umaphip_knn_index *index);

/* ------------------------------------------------------------------------
 * Host graph and initialisation steps (host arrays, multi-threaded).
 * ------------------------------------------------------------------------ */
/* Fuzzy simplicial set. Replaces StandardUmapOps::fuzzy_graph
 * (crates/umap/src/backend.rs:89-111; kernels/standard_umap/graph.rs:
 * 21-564): rho/sigma per row, directed memberships [n,k], and the
 * symmetrised (or bipartite) COO of capacity 2*n*k grouped by head and
 * sorted by tail; *used_out = entries written. NUMERICAL (with the
 * -2/-2/-1 marker in slot 0) on invalid neighbour ids. */
umaphip_status umaphip_fuzzy_graph(const int32_t *indices, const float *distances, int64_t n,
                                   int64_t k, float local_connectivity, float bandwidth, float mix,
                                   int32_t bipartite, int64_t n_reference, int32_t *heads,
                                   int32_t *tails, float *weights, float *rho, float *sigma,
                                   float *directed, int64_t *used_out);
/* Positive-edge pruning (weights below max / prune_divisor become zero;
 * prune_divisor <= 0 disables it; estimator.rs:2753-2798) then compact_graph
 * (crates/umap-core/src/modern_layout.rs:303-430): CSR offsets [rows+1],
 * column-sorted indices, merged weights, f64 periods and initial clocks
 * (capacity n_edges each); *n_compact = E. */
umaphip_status umaphip_compact_graph(int64_t rows, int64_t columns, int64_t n_edges,
                                     const int32_t *heads, const int32_t *tails,
                                     const float *weights_in, double prune_divisor, int32_t epochs,
                                     float rate, int32_t transform, uint32_t *offsets,
                                     uint32_t *indices, float *weights, double *periods,
                                     double *negative_periods, double *next_positive,
                                     double *next_negative, int64_t *n_compact);
/* fit_curve (crates/umap-core/src/config.rs:164-330). */
umaphip_status umaphip_fit_curve(float min_dist, float spread, float *a, float *b);
/* Random U(-10,10) initialisation from StdRng(seed) then centring and
 * isotropic (mode 0) or per-axis (mode 1) robust rescaling
 * (crates/umap-core/src/init.rs:1519-1530, init_rescale.rs:77-203); mode -1
 * returns the raw draws (for umaphip_init_rescale with mode 2). */
umaphip_status umaphip_random_init(int64_t n, int64_t d, uint64_t seed, float repulsion,
                                   float scale_factor, int32_t mode, float *out);
/* init_rescale::rescale (crates/umap-core/src/init_rescale.rs:77-203) in
 * place on values [n, d]: centre each axis, then scale isotropically
 * (mode 0), per axis (mode 1), or so that the weight-median edge length of
 * the graph is scale_factor (mode 2, graph_edge; the graph is a CSR with
 * offsets [n + 1] and its edges in CSR order; ignored by modes 0 and 1). */
umaphip_status umaphip_init_rescale(float *values, int64_t n, int64_t d, const int64_t *offsets,
                                    const int32_t *indices, const float *weights, int32_t mode,
                                    float repulsion, float scale_factor);
/* Transform initial placement from query neighbours and their directed
 * memberships (crates/umap/src/standard_umap/estimator.rs:2385-2481).
 * embedding is the fitted [n, d]; out is [q, d]. */
umaphip_status umaphip_transform_init(const int32_t *indices, const float *directed, int64_t q,
                                      int64_t k, const float *embedding, int64_t n, int64_t d,
                                      float *out);
/* The fit graph on the device: fuzzy_graph (symmetric fuzzy union,
 * kernels/standard_umap/graph.rs:21-433) followed by pruning below
 * max / prune_divisor (estimator.rs:531-560) and compact_graph
 * (crates/umap-core/src/modern_layout.rs:188-208, 303-430), in the host
 * steps' operation order with the device's expf. Device pointers: indices
 * and distances [n, k] (self first, as the fit's kNN gives them); outputs
 * offsets [n + 1] and indices/weights/periods/clocks with capacity 2 n k.
 * *n_edges (host) receives the edge count. Synchronises the stream. */
umaphip_status umaphip_fit_graph(const int32_t *indices, const float *distances, int64_t n,
                                 int64_t k, float local_connectivity, float bandwidth, float mix,
                                 int32_t epochs, float rate, double prune_divisor, uint32_t *offsets,
                                 uint32_t *out_indices, float *weights, double *periods,
                                 double *negative_periods, double *next_positive,
                                 double *next_negative, int64_t *n_edges, void *stream);
/* The transform front-end on the device, for one request's q query rows:
 * bipartite fuzzy memberships (fuzzy_graph with bipartite = true,
 * kernels/standard_umap/graph.rs:117-433), the initial placement
 * (estimator.rs:2385-2481), pruning below max / epochs (estimator.rs:
 * 2753-2798) and compact_graph (crates/umap-core/src/modern_layout.rs:
 * 188-208, 303-430), in the host steps' operation order with the device's
 * expf. Device pointers: indices/distances [q, k] from the kNN, the fitted
 * embedding [n, d]; outputs offsets [q + 1], indices/periods/clocks with
 * capacity q * k, init [q, d]. *n_edges (host) receives the kept edges.
 * local_connectivity is the transform's value (max(fit value - 1, 0)).
 * Synchronises the stream. The transform server builds its graphs with the
 * same kernels, so served and single transforms agree bit for bit. */
umaphip_status umaphip_transform_graph(const int32_t *indices, const float *distances, int64_t q,
                                       int64_t k, const float *embedding, int64_t n, int64_t d,
                                       float local_connectivity, int32_t epochs, float rate,
                                       uint32_t *offsets, uint32_t *out_indices, double *periods,
                                       double *negative_periods, double *next_positive,
                                       double *next_negative, float *init, int64_t *n_edges,
                                       void *stream);

/* ------------------------------------------------------------------------
 * Recursive (multilevel) initialisation, init="recursive" or "auto"
 * (crates/umap-core/src/recursive_init/mod.rs:48-320 with the coarse
 * layouts of crates/umap/src/standard_umap/estimator.rs:2506-2589).
 * Host arrays; the coarse layouts run on the current device.
 * ------------------------------------------------------------------------ */
/* RecursiveInitConfig (crates/umap-core/src/config.rs:382-430). */
typedef struct {
    int64_t coarsening_ratio;       /* 2, 3 or 4 */
    int64_t base_init_threshold;    /* >= 2 */
    int64_t n_embedding_epochs;     /* >= 1, epochs of the finest coarse layout */
    int32_t hub_selection;          /* 0 degree, 1 diverse, 2 random */
    int64_t anchor_sample_size;     /* >= 1 */
} umaphip_recursive_params;

/* LevelParams (crates/umap-core/src/recursive_init/mod.rs:494-561). */
typedef struct {
    int64_t parts;                  /* requested coarse parts */
    int64_t epochs;                 /* coarse layout epochs */
    float a, b;                     /* the curve to the 1/4 power per level */
    int32_t good_initialization;
    double curve_a, curve_b;
} umaphip_recursive_level;

typedef struct {
    int32_t depth;                  /* levels visited, the base case included */
    int64_t rows[32];               /* graph rows at depth 1, 2, ... */
    int64_t edges[32];
    int32_t sweeps[32];             /* label propagation sweeps (0 at the base) */
    double host_ms;                 /* host graph and geometry steps */
    double layout_ms;               /* coarse layouts, wall clock */
    double stage_ms[8];             /* host time by step, summed over levels: anchor,
                                       hubs, propagation, outliers and remap,
                                       aggregation, transition, prolongation and
                                       alignment, base case */
} umaphip_recursive_stats;

/* recursive_init::run (mod.rs:48-123): coordinates [n, d] before the final
 * rescale (apply umaphip_init_rescale, mode as configured, for
 * recursive_init::finish, mod.rs:27-45). The graph is the pruned positive
 * fit graph as a square CSR with sorted columns and symmetric support
 * (offsets [n + 1], indices and weights [offsets[n]]); data is the [n,
 * features] training matrix; (a, b) the fitted curve; seed the fit's seed.
 * INVALID_INPUT for a malformed or asymmetric graph, INVALID_CONFIG for
 * out-of-range parameters. stats may be NULL. */
umaphip_status umaphip_recursive_init(int64_t n, const int64_t *offsets, const int32_t *indices,
                                      const float *weights, const float *data, int64_t features,
                                      int64_t d, float a, float b,
                                      const umaphip_recursive_params *params, uint64_t seed,
                                      float *out, umaphip_recursive_stats *stats);
/* The steps of the recursion, for tests and for callers that drive their
 * own levels (each mirrors the public function named). */
/* LevelParams::first (previous == NULL, curve (a, b)) or previous.next(n)
 * (mod.rs:510-539); *has_level = 0 when n is a base case. */
umaphip_status umaphip_recursive_level_params(int64_t n, const umaphip_recursive_level *previous,
                                       float a, float b, const umaphip_recursive_params *params,
                                       umaphip_recursive_level *out, int32_t *has_level);
/* initialize_labels (graph.rs:140-197); labels [n], -1 for non-hubs. */
umaphip_status umaphip_recursive_hubs(int64_t n, const int64_t *offsets, const int32_t *indices,
                                      const float *weights, int64_t parts, int32_t method,
                                      uint64_t seed, int32_t *labels);
/* propagate_labels (graph.rs:199-253); labels in/out. */
umaphip_status umaphip_recursive_propagate(int64_t n, const int64_t *offsets,
                                           const int32_t *indices, const float *weights,
                                           int32_t share, int32_t *labels, int32_t *sweeps);
/* label_outliers (graph.rs:255-297); labels in/out. */
umaphip_status umaphip_recursive_outliers(int64_t n, const int64_t *offsets,
                                          const int32_t *indices, uint64_t seed,
                                          int32_t *labels);
/* remap_labels (graph.rs:299-314); labels in/out, *parts = part count. */
umaphip_status umaphip_recursive_remap(int32_t *labels, int64_t n, int64_t *parts);
/* aggregate_graph (graph.rs:380-460): summed coarse CSR, outputs
 * out_offsets [parts + 1] and capacity offsets[n] for indices/weights. */
umaphip_status umaphip_recursive_aggregate(int64_t n, const int64_t *offsets,
                                           const int32_t *indices, const double *weights,
                                           const int32_t *labels, int64_t parts,
                                           int64_t *out_offsets, int32_t *out_indices,
                                           double *out_weights);
/* transition_graph (graph.rs:462-533): outputs out_offsets [n + 1] and
 * capacity 2 * offsets[n] for indices/weights. */
umaphip_status umaphip_recursive_transition(int64_t n, const int64_t *offsets,
                                            const int32_t *indices, const double *weights,
                                            int64_t *out_offsets, int32_t *out_indices,
                                            float *out_weights);
/* anchor_reference (geometry.rs:266-427) with the run's generator
 * (StdRng(stream_seed(seed, 1, 0))): values [n, d], mask [n] (1 = sampled). */
umaphip_status umaphip_recursive_anchor(const float *data, int64_t n, int64_t features,
                                        int64_t d, int64_t sample_size, uint64_t seed,
                                        float *values, uint8_t *mask);
/* base_case (geometry.rs:460-539). */
umaphip_status umaphip_recursive_base_case(int64_t n, const int64_t *offsets,
                                           const int32_t *indices, const float *weights,
                                           const float *anchor, const uint8_t *mask, int64_t d,
                                           int64_t depth, float *out);
/* prolongate (geometry.rs:541-582); coarse [parts, d]. */
umaphip_status umaphip_recursive_prolongate(int64_t n, const int64_t *offsets,
                                            const int32_t *indices, const float *weights,
                                            const int32_t *labels, const float *coarse,
                                            int64_t parts, int64_t d, float *out);
/* procrustes_align (geometry.rs:584-690); layout [n, d] in place. */
umaphip_status umaphip_recursive_procrustes(float *layout, int64_t n, const float *anchor,
                                            const uint8_t *mask, int64_t d);

/* ------------------------------------------------------------------------
 * Compatibility (classic) SGD layout, optimizer="compatibility".
 * ------------------------------------------------------------------------ */
/* One epoch. Replaces StandardUmapOps::optimize_epoch
 * (crates/umap/src/backend.rs:234-262; CubeCL body
 * kernels/standard_umap/layout.rs:795-1060). Device pointers, contiguous
 * [rows, d] coordinates. Clocks and rng_state [h,3] are advanced in place;
 * head is updated in place with atomic adds (fit also moves attraction
 * tails, which are the same buffer). head_offsets/head_edges list each head
 * row's positive-weight edge ids in ascending order. mask is an
 * n_edges-entry i32 workspace. move_other must equal shared, as in the
 * reference. The coordinates are not unique (concurrent updates); the clocks
 * and the generator states are. */
typedef struct {
    int64_t h, n, d, n_edges;
    float *head;                       /* [h,d] in/out */
    const float *tail;                 /* [n,d] fixed tail (transform); ignored when shared */
    const int32_t *tails;              /* [n_edges] */
    const float *weights;              /* [n_edges] (zero: never fires) */
    const double *epochs_per_sample;
    double *epoch_of_next_sample;
    const double *epochs_per_negative_sample;
    double *epoch_of_next_negative_sample;
    int64_t *rng_state;                /* [h,3] */
    const int32_t *head_offsets;       /* h + 1 */
    const int32_t *head_edges;
    int32_t *mask;                     /* workspace [n_edges] */
    uint32_t epoch;
    float a, b, gamma, alpha, negative_sample_rate;
    int32_t move_other, shared;
} umaphip_classic_args;
umaphip_status umaphip_classic_epoch(const umaphip_classic_args *args, void *stream);

/* Whole compatibility layout from host arrays (allocates device memory).
 * Replaces estimator::optimize (crates/umap/src/standard_umap/estimator.rs:
 * 2868-2982) with its schedule (prepare_schedule :2753-2798: prune below
 * max / (n_epochs > 10 ? n_epochs : default_epochs), f64 intervals
 * max / w, negative intervals / max(rate, 1)), head CSR (:3080-3131),
 * generator seeding from StdRng(seed) and each row's first coordinate
 * (:2800-2829) and the linear learning-rate decay. heads/tails/weights are
 * the fuzzy graph's COO. The optional outputs receive the final clocks (E)
 * and generator states (3h). */
umaphip_status umaphip_classic_optimize(int64_t h, int64_t n, int64_t d, float *head,
                                        const float *tail, const int32_t *heads,
                                        const int32_t *tails, const float *weights,
                                        int64_t n_edges, int32_t n_epochs, int32_t default_epochs,
                                        float a, float b, float learning_rate, float gamma,
                                        int32_t negative_sample_rate, uint64_t seed,
                                        int32_t shared, double *next_positive_out,
                                        double *next_negative_out, int64_t *rng_out);

/* ------------------------------------------------------------------------
 * Continuous-batching transform server (no reference counterpart: it serves
 * FittedUMAP::transform, crates/umap/src/standard_umap/estimator.rs:
 * 1367-2264, for many concurrent requests). The fitted references, the
 * embedding and the per-epoch projection tables of the embedding stay
 * resident; requests are admitted while a fixed device arena has room, their
 * kNN runs in batches, and each batch's transform layouts (all epochs of all
 * its requests) run in one launch. A request's result equals what
 * UMAP.transform computes for it alone, bit for bit (same tile sweep and
 * per-row summation order).
 * ------------------------------------------------------------------------ */
typedef struct umaphip_server umaphip_server;
typedef struct {
    int32_t n_neighbors;            /* k = min(n_neighbors, n) */
    int32_t epochs;                 /* transform epochs */
    float a, b;                     /* fitted curve */
    float learning_rate;            /* fit learning rate; the transform uses / 4 */
    float repulsion;                /* gamma */
    float local_connectivity;       /* fit value; the transform uses max(lc - 1, 0) */
    float set_op_mix_ratio;
    int64_t negative_sample_rate;
    int64_t negative_selection_range;   /* must be below n (projection windows) */
    int32_t has_scale;
    float scale;
    int64_t adaptation_samples;     /* only selects the default scale formula */
    int32_t exclude_graph_neighbors;
    uint64_t transform_seed;
    int64_t max_batch_rows;         /* rows per kNN batch and per request */
    int64_t memory_bytes;           /* admission arena; 0 = min(free/4, 1 GiB) */
} umaphip_server_params;
typedef struct {
    int64_t requests, rows;
    int64_t steps;                  /* layout launches (one per kNN batch with a layout) */
    int64_t knn_batches;
    double mean_active;             /* requests per layout launch */
    int64_t peak_active;            /* most requests in flight after their kNN */
    int64_t admission_waits, memory_budget, memory_in_use;
    /* mean per-request stage times: waiting for admission, kNN batch (admit
     * to kNN seen finished), host graph + upload + launch, layout + retire */
    double mean_queue_ms, mean_knn_ms, mean_graph_ms, mean_layout_ms;
} umaphip_server_counters;
/* Builds the resident model (copies reference [n, features] and embedding
 * [n, d] from host memory) and starts the server thread. */
umaphip_status umaphip_server_create(const float *reference, int64_t n, int64_t features,
                                     const float *embedding, int64_t d,
                                     const umaphip_server_params *params, umaphip_server **out);
/* Queues q rows (host [q, features], copied); returns a ticket. */
umaphip_status umaphip_server_submit(umaphip_server *server, const float *queries, int64_t q,
                                     int64_t *ticket);
/* Waits up to timeout_s (0: poll) for a ticket. When *done = 1 the result
 * [q, d] is in out, latency_ms is submit-to-completion, and the ticket is
 * released; a failed request returns its status. */
umaphip_status umaphip_server_wait(umaphip_server *server, int64_t ticket, double timeout_s,
                                   float *out, double *latency_ms, int32_t *done);
umaphip_status umaphip_server_stats(umaphip_server *server, umaphip_server_counters *stats);
void umaphip_server_destroy(umaphip_server *server);

/* ------------------------------------------------------------------------
 * Batched layouts (no reference counterpart: many independent
 * optimize_compact fits, crates/umap/src/standard_umap/modern_runtime.rs:
 * 1011-1378, in one launch per epoch). Every graph is an Adam fit in global
 * negative mode (rows <= negative_selection_range) with its own seed; the
 * per-epoch orders are generated up front and the whole epoch loop is
 * captured once as a hipGraph. Graph g's arrays are concatenated: offsets
 * holds rows[g] + 1 local entries per graph, indices/clocks n_edges[g]
 * local entries, head rows[g] * d floats.
 * ------------------------------------------------------------------------ */
typedef struct umaphip_batch umaphip_batch;
umaphip_status umaphip_batch_create(int32_t graphs, const int64_t *rows, const int64_t *n_edges,
                                    const uint32_t *offsets, const uint32_t *indices,
                                    const double *periods, const double *next_positive,
                                    const double *negative_periods, const double *next_negative,
                                    const float *head, int64_t d,
                                    const umaphip_layout_params *params, const uint64_t *seeds,
                                    umaphip_batch **out);
/* Autotuner: times every tile shape (0: 64 rows x 256 lanes, 1: 32 x 128,
 * 2: 16 x 64 -- one wave per tile) over the first probe_epochs epochs and
 * keeps the fastest; ms_out (optional) receives the three timings. */
umaphip_status umaphip_batch_tune(umaphip_batch *batch, int32_t probe_epochs, int32_t *shape_out,
                                  double *ms_out);
umaphip_status umaphip_batch_set_shape(umaphip_batch *batch, int32_t shape);
/* Launch form of later runs and tuning (no reference counterpart; the
 * reference runs one layout at a time): persistent = 1 runs every epoch of
 * every graph in one cooperative launch (issued directly, whatever
 * use_graph says), the workgroups meeting at a grid barrier between epochs;
 * 0 launches one segment sweep per epoch. Both give the same bits. */
umaphip_status umaphip_batch_set_persistent(umaphip_batch *batch, int32_t persistent);
/* Clock schedule of the batch (no reference counterpart; the reference
 * advances both clocks inside its gradient kernel, kernels/standard_umap/
 * modern_layout.rs:166-253): enabled = 1 (the default when allowed) runs
 * every edge's clocks through all epochs in one pass before the first epoch
 * and stores its term count per epoch (one byte per edge and epoch), which
 * the tile sweeps read instead of the four f64 clock words; 0 advances the
 * clocks inside the sweeps. Both give the same bits. A batch allows it when
 * those bytes fit 8 GiB, every count provably fits a byte and the memory can
 * be allocated; otherwise it stays off. umaphip_batch_clock_schedule returns the epochs the schedule
 * covers (0: off). */
umaphip_status umaphip_batch_set_clock_schedule(umaphip_batch *batch, int32_t enabled);
int32_t umaphip_batch_clock_schedule(const umaphip_batch *batch);
/* Resets to the uploaded initial state and enqueues every epoch: as one
 * hipGraph launch (use_graph = 1, captured on first use) or as plain
 * launches. Does not synchronise. */
umaphip_status umaphip_batch_run(umaphip_batch *batch, int32_t use_graph);
/* Synchronises; copies the concatenated head (and optionally clocks) to
 * host; device_ms = duration of the last run's epochs. */
umaphip_status umaphip_batch_download(umaphip_batch *batch, float *head, double *next_positive,
                                      double *next_negative, double *device_ms);
void umaphip_batch_destroy(umaphip_batch *batch);

#ifdef __cplusplus
}
#endif
#endif
