// Internal launch interface of the modern-layout kernels (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace umaphip {

// Generic per-slot gradient (any d, strided coordinates, reference slot
// order). Mirrors gradient_kernel (kernels/standard_umap/modern_layout.rs:
// 128-257) lane-for-slot.
struct RowwiseArgs {
    int64_t h, n, d;
    const float *head; int64_t hs0, hs1;
    const float *tail; int64_t ts0, ts1;
    const int32_t *offsets, *indices, *source_order, *negative_order, *source_ranks;
    const double *periods; double *next_positive;
    const double *negative_periods; double *next_negative;
    float *updates; int64_t us0, us1;
    uint32_t epoch; float a, b, gamma, scale; uint32_t range;
    float multiplier;
    int shared, negatives, exclude;
    const float *scale_ptr;   // device scale (may be null: use `scale`)
};
void launch_gradient_rowwise(const RowwiseArgs &a, int64_t start, int64_t end, hipStream_t s);

// Lazy Adam over [h, d] strided buffers (adam_kernel, :495-523).
void launch_adam(int64_t h, int64_t d, float *head, int64_t hs0, int64_t hs1, float *upd,
                 int64_t us0, int64_t us1, float *m, int64_t ms0, int64_t ms1, float *v,
                 int64_t vs0, int64_t vs1, float alpha, float c1, float c2, hipStream_t s);
// Momentum apply over source_order slots [start, end) (apply_kernel, :525-543).
void launch_momentum_apply(int64_t d, float *head, int64_t hs0, int64_t hs1, const float *vel,
                           int64_t vs0, int64_t vs1, const int32_t *order, int64_t start,
                           int64_t end, hipStream_t s);
// velocity *= mu (decay_kernel, :545-554).
void launch_momentum_decay(int64_t h, int64_t d, float *vel, int64_t vs0, int64_t vs1, float mu,
                           hipStream_t s);

// Tiled edge sweep for d = 2 with contiguous float2 coordinates.
struct SweepArgs {
    int64_t h, n;
    const float *pos;         // [h,2] Jacobi snapshot of the source rows
    float *pos_out;           // fused Adam output [h,2]
    const float *tgt;         // attraction targets: pos (fit) or fixed tail (transform)
    const float *table;       // negatives by rank: table[r] = position of order[r]; may be null
    const int32_t *order;     // negative order (ids by rank)
    const int32_t *ranks;     // [h] window ranks (large windowed layouts)
    const int32_t *offsets, *indices;
    const double *periods; double *next_positive;
    const double *negative_periods; double *next_negative;
    float *updates;           // ACCUM mode: updates[i] += multiplier * g
    float *m, *v;             // fused Adam moments
    const int32_t *block_of_row; int32_t block_id;  // momentum block filter (may be null)
    const float *scale_ptr;   // device scale (may be null: use `scale`)
    float scale;
    uint32_t epoch; float a, b, gamma; uint32_t range;
    float multiplier, alpha, c1, c2;
    int shared, negatives, exclude, fused;
    // tile sweep: this epoch's clock-schedule slab ([h rows' edges] term
    // counts, launch_edge_schedule), or null to advance the clocks here
    const uint8_t *sched;
};
void launch_edge_sweep(const SweepArgs &a, hipStream_t s);

// Batched tile sweep: segment s owns global tiles [tile0, tile0 + ceil(h /
// tile_rows())); segments are in ascending tile0 order.
struct SweepSegment {
    SweepArgs args;
    int64_t tile0;
};
// Tile shapes: 0 = 64 rows x 256 lanes, 1 = 32 x 128, 2 = 16 x 64.
constexpr int kTileShapes = 3;
int64_t tile_rows(int shape = 0);
void launch_segment_sweep(const SweepSegment *segs, int32_t n_segs, int64_t total_tiles,
                          hipStream_t s, int shape = 0);
// Every epoch of a batch in one cooperative launch: segs holds `epochs`
// consecutive tables of n_segs segments (epoch t's at t * n_segs); a grid
// barrier separates the epochs; counters: 8 * epochs ints of workspace
// (cleared by the launcher). Results equal one launch_segment_sweep per
// epoch, bit for bit. Returns the cooperative launch's status.
// Negative tables of every graph of a batch for epoch t (2-D): table[i] =
// pos[r0 + order[order_start[g] + t * h + (i - r0)]] for row i of graph g
// (rows r0 .. r0 + h - 1, graph_of_row[i] = g).
void launch_batch_table(const float *pos, const int32_t *orders, const int64_t *row_start,
                        const int64_t *order_start, const int32_t *graph_of_row,
                        int64_t total_rows, int32_t t, float *table, hipStream_t s);
hipError_t launch_persistent_segments(const SweepSegment *segs, int32_t n_segs,
                                      int64_t total_tiles, int32_t epochs, int32_t *counters,
                                      hipStream_t s, int shape = 0);

// Whole transform layouts in one launch. In a transform every row moves
// alone against the fixed tail, and the tail's negative tables for each
// epoch are fixed in advance, so a workgroup can run all epochs of its
// tile without any grid-wide step. Segment s holds one request: `args`
// carries its epoch-invariant arguments with pos = the initial positions
// and pos_out = a second buffer (epoch t reads buffer t & 1); tiles of
// transform_tile_rows() rows; segments in ascending tile0 order.
struct TransformSegment {
    SweepArgs args;
    int64_t tile0;
    int32_t *ranks;                   // [h] window ranks (h > 1024), rewritten each epoch
    unsigned long long *status;       // min(epoch << 32 | row) over non-finite projections
};
struct TransformTables {
    const float *alpha, *c1, *c2;     // [epochs] schedule and bias corrections
    const int32_t *iota;              // identity order (h <= 1024, epoch 0)
    const float *tail;                // [n, 2]
    const int32_t *order;             // [epochs, n] projection order of each direction
    const float *sorted;              // [epochs, n] sorted projections
    const float *table;               // [epochs, n, 2] tail positions by rank
    const float *dirs;                // [epochs, 2]
    int64_t n;
    int32_t epochs;
};
int64_t transform_tile_rows();
void launch_transform_layouts(const TransformSegment *segs, int32_t n_segs, int64_t total_tiles,
                              const TransformTables &tables, hipStream_t s);

// Edge-balanced unit sweep for Adam layouts with H > 1024 and d = 2: one
// wave per unit_edges() consecutive CSR edges; writes per-row gradients
// (grad) and, for rows crossing a unit boundary, per-unit fragments.
struct UnitSweepArgs {
    SweepArgs s;               // positions, negatives and parameters (pos_out, m, v unused)
    double *rec;               // [E] records {next_positive, next_negative, period, j | e mod M << 32},
                               // compact: {next_positive, next_negative, j | samples (f32 bits) << 32}
    uint16_t *fire;            // [E] firing shadow
    const int32_t *unit_k;     // [units][2] first/last nonzero-degree row index of each unit
    const int32_t *unit_frag;  // [units][2] fragment slot of the unit's first/last row, or -1
    const int32_t *nz_rows;    // nonzero-degree row ids (null when every row has an edge)
    const int32_t *nz_off;     // CSR start of each nonzero-degree row, then E
    float *grad, *frag;
    // rank-ordered sweep: [n] set to 1 when the row's gradient is split over
    // units (grad then holds the bits of its first and last unit); zero
    // otherwise (the Adam step clears it after reading it)
    uint8_t *split;
    const int32_t *pref;       // rank-ordered sweep: [n + 1] sequence start of each rank, or null
    const int32_t *rank_seg;   // rank-ordered sweep: [n] CSR start of order[r]
    int64_t units, n_edges;
    int64_t xcd_chunk;         // units per XCD slice (set by the launcher)
    int32_t row_cap;           // unit_narrow_rows() when no unit holds more rows, else unit_edges()
    double rate;               // negative_sample_rate
    uint64_t t1;               // (epoch + 1) mod M (set by the launcher)
    double inv_mod;            // 1 / M (set by the launcher)
    bool compact;              // 24-byte records (rank-ordered sweep only)
    double per_num;            // compact: period = per_num / samples
    // this epoch's slab of the clock schedule ([E] term counts, 0 when the
    // edge does not fire), or null to advance the clocks from the records
    const uint8_t *sched;
    // Epoch trace (null tr_count: off). Per CSR edge: tr_count[e] = the
    // negatives drawn this epoch (left untouched when the edge does not
    // fire), tr_cand[e * tr_cap + p] = candidate id of sample p and
    // tr_kept[...] = 1 unless exclusion dropped it; tr_overflow is set when
    // an edge draws more than tr_cap; tr_paths[0] / [1] count the units that
    // took the term-parallel / per-edge path.
    int32_t *tr_count, *tr_cand;
    uint8_t *tr_kept;
    int32_t tr_cap;
    int32_t *tr_overflow, *tr_paths;
};
int64_t unit_edges();
int64_t unit_narrow_rows();
void launch_unit_sweep(UnitSweepArgs a, hipStream_t s);
void launch_unit_pack(const double *npos, const double *nneg, const double *periods,
                      const int32_t *indices, int64_t E, uint64_t modulus, bool compact,
                      double per_num, double *rec,
                      uint16_t *fire, hipStream_t s);
void launch_unit_unpack(const double *rec, int64_t E, bool compact, double *npos, double *nneg,
                        hipStream_t s);
// Clock schedule of epochs [t0, t0 + span) for the unit sweep: slab
// t - t0 (at sched + (t - t0) * stride) holds each edge's term count in epoch
// t (1 + negatives, 0 when it does not fire); the records advance to the
// state after epoch t0 + span - 1. A count above 255 ORs 4 into *flag.
int64_t unit_schedule_stride(int64_t E);
// The same schedule over clocks kept as four f64 arrays (npos and nneg are
// advanced in place), for the tile sweep's SweepArgs::sched.
void launch_edge_schedule(double *npos, double *nneg, const double *periods,
                          const double *negative_periods, int64_t E, bool negatives, int32_t t0,
                          int32_t span, uint8_t *sched, int64_t stride, int32_t *flag,
                          hipStream_t s);
void launch_unit_schedule(double *rec, int64_t E, bool compact, double per_num, double rate,
                          bool negatives, int32_t t0, int32_t span, uint8_t *sched,
                          int64_t stride, int32_t *flag, hipStream_t s);
// Whether every period equals per_num / float(per_num / period) exactly, so
// that 24-byte records (f32 samples instead of the f64 period) reproduce it.
bool periods_compactable(const double *periods, int64_t E, double per_num);
// Rank-ordered sweep helpers: from rank_deg [n] (the degree of order[r],
// written with rank_seg by the projection sort's last pass), pref[r] = edges
// of the ranks below r (block_sums: rank_prefix_blocks(n) ints) and unit_k
// [units][2], the ranks holding each unit's first and last slot; and the
// matching Adam pass.
int64_t rank_prefix_blocks(int64_t n);
void launch_rank_prefix(const int32_t *rank_deg, int64_t n, int32_t total_edges, int32_t *pref,
                        int32_t *block_sums, int32_t *unit_k, hipStream_t s);
// grad_out (may be null) receives each row's assembled gradient, the value
// the Adam step consumes. With rows_out (4 floats per row), it also writes
// each row's packed record {x, y, CSR start, degree} (launch_pack_rows).
void launch_unit_adam_ranked(int64_t h, const float *pos, float *pos_out, const float *grad,
                             uint8_t *split, const float *frag, float *m, float *v,
                             float alpha, float c1, float c2, hipStream_t s,
                             float *grad_out = nullptr, const int32_t *csr_off = nullptr,
                             float *rows_out = nullptr);
// rows[i] = {x_i, y_i, CSR start, degree} (the two integers as float bits)
// for the projection sort's last pass: one 16-byte gather per key instead of
// a position and two offsets.
void launch_pack_rows(const float *pos, const int32_t *csr_off, int64_t h, float *rows,
                      hipStream_t s);
void launch_unit_adam(int64_t h, const float *pos, float *pos_out, const float *grad,
                      const int32_t *row_split, const int32_t *split_off, const float *frag,
                      float *m, float *v, float alpha, float c1, float c2, hipStream_t s,
                      float *grad_out = nullptr);

// FP64 fuzzy-set cross entropy over the positive edges (doubles of workspace:
// cross_entropy_workspace(h)); *out is a device double.
int64_t cross_entropy_workspace(int64_t h);
void launch_cross_entropy(const float *head, int64_t h, const float *tail, int64_t d,
                          const int32_t *offsets, const int32_t *indices, const float *weights,
                          float a, float b, double *workspace, double *out, hipStream_t s);

// Debug: per-edge negative candidates (same selection code).
void launch_debug_negatives(const RowwiseArgs &a, int32_t cap, int32_t *count_out,
                            int32_t *cand_out, uint8_t *kept_out, int32_t *overflow,
                            hipStream_t s);

// Projection helpers.
void launch_project(const float *x, int64_t rows, int64_t d, int64_t s0, int64_t s1,
                    const float *dir, float *out, hipStream_t s);
// keys[i] = radix key of the projection (d = 2 contiguous or generic) and
// non-finite tracking: status = min(tag | row) over non-finite rows.
void launch_project_keys(const float *x, int64_t rows, int64_t d, const float *dir,
                         uint32_t *keys, float *proj_out, unsigned long long *status,
                         unsigned long long tag, hipStream_t s);
void launch_keys_from_values(const float *values, int64_t n, uint32_t *keys,
                             unsigned long long *status, hipStream_t s);
void launch_iota(int32_t *out, int64_t n, hipStream_t s);
size_t sort_pairs_workspace(int64_t n);
// Stable ascending LSD radix sort of (key, value) pairs; vals_in may be null
// (values = input positions). The _ranked form also writes, from its last
// pass, ranks[value] = position and (table_src non-null, d = 2)
// table_out[position] = table_src[value].
void launch_sort_pairs(void *ws, size_t ws_bytes, const uint32_t *keys_in, uint32_t *keys_out,
                       const int32_t *vals_in, int32_t *vals_out, int64_t n, hipStream_t s);
void launch_sort_pairs_ranked(void *ws, size_t ws_bytes, const uint32_t *keys_in,
                              uint32_t *keys_out, const int32_t *vals_in, int32_t *vals_out,
                              int64_t n, int32_t *ranks, const float *table_src, float *table_out,
                              hipStream_t s);
// Fit refresh for d = 2: keys of x . dir (non-finite rows tracked in status as
// launch_project_keys does), then the stable sort of (key, row), one launch
// per 8-bit pass, with ranks[row] = position and table_out[position] = x[row].
// With csr_off (the CSR offsets of the n rows), the last pass also writes
// seg_out[position] / deg_out[position], the CSR start and degree of the row.
// With rows (launch_pack_rows of x and csr_off), the last pass reads the
// position, CSR start and degree from that one record; ranks may be null.
// A pass whose wait for earlier tiles exceeds its spin limit sets bit 2 of
// *err. The workspace (project_sort_workspace bytes, n < 2^31) needs its
// digit totals clear on entry (clear_project_sort_workspace; every sort
// leaves them clear).
size_t project_sort_workspace(int64_t n);
void clear_project_sort_workspace(void *ws, size_t ws_bytes, int64_t n, hipStream_t s);
void launch_project_sort_ranked(void *ws, size_t ws_bytes, const float *x, int64_t n,
                                const float *dir, unsigned long long *status,
                                unsigned long long tag, uint32_t *keys, uint32_t *keys_out,
                                int32_t *vals_out, int32_t *ranks, float *table_out, hipStream_t s,
                                const int32_t *csr_off, int32_t *seg_out, int32_t *deg_out,
                                int32_t *err, const float *rows = nullptr);
void launch_values_from_keys(const uint32_t *keys, int64_t n, float *values, hipStream_t s);
void launch_inverse_ranks(const int32_t *order, int64_t n, int32_t *ranks, hipStream_t s);
void launch_lower_bound(const float *sorted, int64_t n, const float *query, int64_t h,
                        int32_t *ranks, hipStream_t s);
// Fit: ranks[order[r]] = r and table[r] = src[order[r]] (d = 2) in one pass.
void launch_rank_and_table(const int32_t *order, int64_t n, const float *src, int32_t *ranks,
                           float *table, hipStream_t s);
void launch_gather_table(const int32_t *order, int64_t n, const float *src, float *table,
                         hipStream_t s);
void launch_block_of_row(const int32_t *source_order, int64_t h, int64_t block,
                         int32_t *block_of_row, hipStream_t s);
void launch_momentum_apply_rows(int64_t h, float *head, const float *vel,
                                const int32_t *block_of_row, int32_t block_id, hipStream_t s);
void launch_status_min_u64(unsigned long long *status, unsigned long long value, hipStream_t s);

// Adaptation.
void launch_resultants(const float *emb, int64_t n, int64_t d, const int32_t *order,
                       const int32_t *offsets, const int32_t *indices, const int32_t *sample_ranks,
                       const int32_t *local_offsets, const int32_t *global_ranks, int64_t samples,
                       uint32_t raw_range, float a, float b, float gamma, int exclude,
                       double *out, hipStream_t s);
// Device-side ratio + anchored scale update; state = {prior, current,
// anchor, has_anchor}, writes scale_out (f32) and an error flag.
constexpr int64_t kDeviceRatioMax = 4096;
void launch_ratio_update(const double *resultants, int64_t samples, double *state,
                         float *scale_out, int32_t *error_flag, hipStream_t s);

}  // namespace umaphip
