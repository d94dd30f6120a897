/*
 * CPU restatement of cuscan's modern UMAP layout path (TEST INFRASTRUCTURE).
 *
 * This library is the correctness oracle and the CPU baseline. It is loaded
 * only by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg. The shipped package (cuscan_amd) never links, imports or calls it.
 *
 * Every function cites the reference file:line it restates. Paths are
 * relative to the cuscan repository root.
 *
 * Parity status of each group is stated in the header comment of
 * umap_oracle.c.
 */
#ifndef UMAP_ORACLE_H
#define UMAP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ RNG */
/* rand 0.10.3 StdRng (ChaCha12, 64-bit block counter, stream 0) seeded by
 * rand_core's SeedableRng::seed_from_u64 (PCG32 expansion). */
typedef struct {
    uint32_t key[8];
    uint64_t block;      /* next block counter to generate */
    uint32_t buf[64];    /* four 16-word blocks */
    int idx;             /* next unread word in buf, 64 = empty */
} orc_rng;

void orc_rng_seed(orc_rng *rng, uint64_t seed);
uint32_t orc_rng_u32(orc_rng *rng);
uint64_t orc_rng_u64(orc_rng *rng);
double orc_rng_f64(orc_rng *rng);
uint32_t orc_rng_range_u32(orc_rng *rng, uint32_t bound);
float orc_rng_range_f32(orc_rng *rng, float low, float high);
void orc_rng_shuffle_i32(orc_rng *rng, int32_t *values, size_t n);
void orc_gaussian_pair(orc_rng *rng, double *first, double *second);
/* ChaCha block function with an arbitrary round count (for test vectors). */
void orc_chacha_block(const uint32_t input[16], int rounds, uint32_t out[16]);

/* ------------------------------------------------------ host building blocks */
void orc_adam_schedule(int epochs, double initial_alpha, int good, double *alpha,
                       double *momentum);
void orc_momentum_schedule(int epochs, double initial_alpha, int good, double *alpha,
                           double *momentum);
uint64_t orc_negative_index(uint64_t edge, uint64_t epoch, uint64_t sample, uint64_t refs);
uint64_t orc_effective_negative_range(uint64_t raw, uint64_t refs);
uint64_t orc_window_negative_index(uint64_t edge, uint64_t epoch, uint64_t sample,
                                   uint64_t rank, uint64_t refs, uint64_t range);
int orc_adaptation_due(int fit, int adam, uint64_t samples, uint64_t raw_range,
                       uint64_t refs, uint64_t epoch, uint64_t total);
float orc_resolved_negative_scale(uint64_t raw_range, uint64_t refs, int has_explicit,
                                  float explicit_scale, int adaptation_enabled);
/* Returns the number of compact edges (>0), or a negative error code.
 * Output arrays must hold `n_edges` entries (offsets: rows + 1). */
long orc_compact_graph(int64_t rows, int64_t columns, int64_t n_edges, const int64_t *row,
                       const int64_t *column, const float *weight, int epochs, float rate,
                       uint32_t *offsets, uint32_t *indices, float *weights, double *periods,
                       double *negative_periods, double *next_positive,
                       double *next_negative);
/* Advance one edge's clocks; returns -1 when it does not fire, else the
 * non-negative negative event count. */
long orc_advance_clock(double *next_positive, double *next_negative, double period,
                       double negative_period, uint64_t epoch);
void orc_accumulate_force(float *gradient, const float *source, const float *target, int d,
                          float a, float b, float gamma, float scale, int small,
                          int negative);

/* ------------------------------------------------------------ the epoch */
typedef struct {
    int64_t h, n, d;
    float *head;  int64_t head_s0, head_s1;
    const float *tail; int64_t tail_s0, tail_s1;   /* ignored when shared */
    const uint32_t *offsets;
    const uint32_t *indices;
    const int32_t *source_order;   /* h entries */
    const int32_t *negative_order; /* n entries */
    const int32_t *source_ranks;   /* h entries (read only for large windows) */
    const double *periods;
    double *next_positive;
    const double *negative_periods;
    double *next_negative;
    float *updates; int64_t upd_s0, upd_s1;
    float *m; float *v;            /* [h, d] contiguous */
    uint32_t epoch;
    float a, b, gamma, scale;
    uint32_t negative_range;
    float alpha, mu, correction1, correction2;
    int momentum, shared, negatives, exclude;
} orc_epoch_args;

/* gradient_kernel over slots [start, end) of source_order, multiplier as given. */
void orc_gradient_block(const orc_epoch_args *args, int64_t start, int64_t end,
                        float multiplier, int threads);
/* gradient_kernel in float64 over the same f32 inputs, written to grad64
 * (h*d, zero-initialised by the caller). Clocks are advanced exactly as the
 * f32 path does. */
void orc_gradient_block64(const orc_epoch_args *args, int64_t start, int64_t end,
                          double multiplier, double *grad64);
/* The same, also adding to magnitude[i] (h entries, zero-initialised by the
 * caller) the Euclidean norm of every term of row i: the row's sum of term
 * magnitudes, which bounds how well any arithmetic that rounds each term
 * can resolve the row's sum (the condition of the sum). */
void orc_gradient_block64_mag(const orc_epoch_args *args, int64_t start, int64_t end,
                              double multiplier, double *grad64, double *magnitude);
void orc_adam_apply(const orc_epoch_args *args, int threads);
void orc_momentum_apply_block(const orc_epoch_args *args, int64_t start, int64_t end);
void orc_momentum_decay(const orc_epoch_args *args);
/* modern_epoch: the composition above. */
void orc_modern_epoch(const orc_epoch_args *args, int threads);
/* Enumerate the negatives the gradient would sample: for every visited
 * firing edge in slot order, writes (edge, candidate, kept) triples. Returns
 * the number of triples (possibly more than capacity, which bounds writes).
 * Clocks are not modified. */
int64_t orc_trace_negatives(const orc_epoch_args *args, int64_t capacity, int64_t *edge_out,
                            int32_t *cand_out, uint8_t *kept_out, int64_t *fire_count);

/* ------------------------------------------------------- projection windows */
void orc_projection_direction(orc_rng *rng, int d, float *direction);
void orc_project(const float *x, int64_t rows, int64_t d, int64_t s0, int64_t s1,
                 const float *direction, float *out);
uint32_t orc_radix_key_f32(float value);
/* Stable ascending argsort; returns -1 or the lowest non-finite row. */
int64_t orc_stable_argsort_f32(const float *values, int64_t n, int32_t *order, float *sorted);
void orc_inverse_ranks(const int32_t *order, int64_t n, int32_t *ranks);
void orc_lower_bound_ranks(const float *sorted, int64_t n, const float *query, int64_t h,
                           int32_t *ranks);

/* ------------------------------------------------------- scale adaptation */
void orc_resultants(const float *emb, int64_t n, int64_t d, const int32_t *order,
                    const uint32_t *offsets, const uint32_t *indices,
                    const int32_t *sample_ranks, const int32_t *local_offsets,
                    const int32_t *global_ranks, int64_t samples, uint32_t raw_range,
                    float a, float b, float gamma, int exclude, double *resultants);
/* Returns the ratio, or a negative number on a numerical error. */
double orc_force_ratio(const double *global, const double *local, int64_t samples);
typedef struct {
    double prior, current, anchor;
    int has_anchor;
} orc_scale_state;
void orc_scale_init(orc_scale_state *s, float scale);
float orc_scale_observe(orc_scale_state *s, double ratio);

/* --------------------------------------------------------- epoch driver */
typedef struct {
    int optimizer;            /* 0 adam, 1 momentum */
    int epochs;
    float a, b, repulsion, initial_alpha;
    int64_t negative_sample_rate;
    int64_t negative_selection_range;
    int has_scale; float scale;
    int64_t adaptation_samples;
    int exclude;
    int good;
    uint64_t seed;
    /* objective trace (the build's FP64 cross entropy, no reference
     * counterpart): with obj_every > 0, after every obj_every-th epoch and
     * after the last one, obj_out[k++] = orc_cross_entropy of the embedding
     * over the positive edges with obj_weights [E] */
    int obj_every;
    const float *obj_weights;
    double *obj_out;
} orc_layout_params;

typedef struct {
    int64_t epochs_run;
    int64_t firing_edges;     /* sum over epochs of F_t */
    int64_t negative_samples; /* sum over epochs of S_t (drawn, incl. rejected) */
    int64_t adaptations;
    int64_t refreshes;
    float final_scale;
} orc_layout_stats;

/* optimize_compact_inner (fit when shared != 0, else transform). head
 * [h,d] is updated in place; tail [n,d] is read for transform. Clock
 * arrays are updated in place. Runs epochs [0, stop_epoch) when
 * stop_epoch > 0 (baseline sampling), else all epochs. Returns 0 or an
 * error code. */
int orc_optimize_compact(int64_t h, int64_t n, int64_t d, float *head, const float *tail,
                         const uint32_t *offsets, const uint32_t *indices,
                         const double *periods, double *next_positive,
                         const double *negative_periods, double *next_negative,
                         const orc_layout_params *params, int shared, int threads,
                         int stop_epoch, orc_layout_stats *stats);

/* ------------------------------------------------------ graph construction */
void orc_exact_knn(const float *query, int64_t nq, const float *reference, int64_t nr,
                   int64_t d, int64_t k, int exclude_self, int64_t query_offset, int threads,
                   int32_t *indices, float *distances);
/* Returns the COO length written (<= 2*n*k) or a negative error. */
long orc_fuzzy_graph(const int32_t *indices, const float *distances, int64_t n, int64_t k,
                     float local_connectivity, float bandwidth, float mix, int bipartite,
                     int64_t n_reference, int32_t *heads, int32_t *tails, float *weights,
                     float *rho, float *sigma, float *directed);
int orc_fit_curve(float min_dist, float spread, float *a, float *b);
int orc_random_init(int64_t n, int64_t d, uint64_t seed, float repulsion, float factor,
                    int mode, float *out);
int orc_transform_init(const int32_t *indices, const float *directed, int64_t q, int64_t k,
                       const float *embedding, int64_t d, float *out);

/* ---- compatibility (classic) SGD: kernels/standard_umap/layout.rs:68-336,
 * estimator.rs:2753-2982. Serial over head vertices in ascending order, each
 * vertex's positive edges in edge-id order. */
int64_t orc_taus_next(int64_t state[3]);  /* advances; returns the signed 32-bit draw */
void orc_classic_epoch(int64_t h, int64_t n, int64_t d, float *head, const float *tail,
                       const int32_t *tails, const float *weights, int64_t E,
                       const double *per, double *npos, const double *nper, double *nneg,
                       int64_t *rng, const int32_t *head_offsets, const int32_t *head_edges,
                       uint32_t epoch, float a, float b, float gamma, float alpha, float rate,
                       int shared);
int orc_classic_optimize(int64_t h, int64_t n, int64_t d, float *head, const float *tail,
                         const int32_t *heads, const int32_t *tails, const float *weights,
                         int64_t E, int32_t n_epochs, int32_t default_epochs, float a, float b,
                         float learning_rate, float gamma, int32_t rate, uint64_t seed,
                         int shared, double *npos_out, double *nneg_out, int64_t *rng_out);

/* FP64 fuzzy-set cross entropy over the positive edges (the build's
 * objective; no reference counterpart), compensated summation in row order. */
double orc_cross_entropy(const float *head, int64_t h, const float *tail, int64_t d,
                         const uint32_t *offsets, const uint32_t *indices, const float *weights,
                         double a, double b);

#ifdef __cplusplus
}
#endif
#endif
