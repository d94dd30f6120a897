// Transform front-end on the device: for each request of a batch, the
// bipartite fuzzy memberships of its query rows against the fitted
// references (smooth kNN: rho, sigma by 64-step bisection), the weighted
// initial placement, positive-edge pruning at max / epochs and the compact
// layout CSR with f64 periods and clocks. This is the chain
// fuzzy_graph(bipartite) -> transform_init -> prepare_schedule pruning ->
// compact_graph of crates/umap/src/standard_umap/estimator.rs:2062-2264,
// 2753-2798, kernels/standard_umap/graph.rs:117-433 and crates/umap-core/
// src/modern_layout.rs:188-208, 303-430, restated per row on the device
// with the host steps' operation order (exp is the device's expf).
//
// Requests are segments of consecutive query rows. Four launches:
//   1. per segment: f64 mean of its distances (the sigma floor);
//   2. per row: rho, sigma, memberships, the row's edges sorted by reference
//      id, the initial position, and the segment's largest weight;
//   3. per segment: kept-edge counts per row and their prefix (offsets);
//   4. per row: indices, periods and clocks of the kept edges.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "common.hpp"
#include "graph_device.hpp"
#include "layout_kernels.hpp"

namespace umaphip {
namespace {

constexpr int kRowThreads = 64;
constexpr int kSegThreads = 256;

__device__ inline int find_segment(const TGraphSegment *segs, int32_t n_segs, int64_t row) {
    int lo = 0, hi = n_segs;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (segs[mid].row0 <= row) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kSegThreads) void tg_mean_kernel(const TGraphSegment *segs,
                                                             const float *dist, int64_t k) {
    __shared__ double s_sum[kSegThreads];
    const TGraphSegment &S = segs[blockIdx.x];
    const int tid = threadIdx.x;
    const int64_t b = S.row0 * k, e = (S.row0 + S.rows) * k;
    double acc = 0.0;
    for (int64_t i = b + tid; i < e; i += kSegThreads) acc += static_cast<double>(dist[i]);
    s_sum[tid] = acc;
    __syncthreads();
    for (int w = kSegThreads / 2; w > 0; w >>= 1) {
        if (tid < w) s_sum[tid] = s_sum[tid] + s_sum[tid + w];
        __syncthreads();
    }
    if (tid == 0) {
        S.state->mean = static_cast<float>(s_sum[0] / static_cast<double>(e - b));
        S.state->wmax_bits = 0u;
        S.state->edges = 0;
        S.state->errors = 0;
    }
}

// The fit's mean distance over [0, m) in two deterministic levels: block b
// sums the b-th contiguous chunk in f64 (fixed tree), then one workgroup adds
// the block sums in order and writes the state.
constexpr int kMeanBlocks = 1024;

__global__ __launch_bounds__(kSegThreads) void fg_mean_partial_kernel(const float *dist, int64_t m,
                                                                      double *partial) {
    __shared__ double s_sum[kSegThreads];
    const int tid = threadIdx.x;
    const int64_t chunk = (m + kMeanBlocks - 1) / kMeanBlocks;
    const int64_t b = static_cast<int64_t>(blockIdx.x) * chunk, e = min<int64_t>(m, b + chunk);
    double acc = 0.0;
    for (int64_t i = b + tid; i < e; i += kSegThreads) acc += static_cast<double>(dist[i]);
    s_sum[tid] = acc;
    __syncthreads();
    for (int w = kSegThreads / 2; w > 0; w >>= 1) {
        if (tid < w) s_sum[tid] = s_sum[tid] + s_sum[tid + w];
        __syncthreads();
    }
    if (tid == 0) partial[blockIdx.x] = s_sum[0];
}

__global__ void fg_mean_final_kernel(const double *partial, int64_t m, TGraphState *state) {
    if (threadIdx.x != 0) return;
    double sum = 0.0;
    for (int b = 0; b < kMeanBlocks; ++b) sum += partial[b];
    state->mean = static_cast<float>(sum / static_cast<double>(m));
    state->wmax_bits = 0u;
    state->edges = 0;
    state->errors = 0;
}

// smooth_row (graph.rs:117-242) on one row's k distances.
__device__ inline void smooth_row(const float *dist, int k, float global_mean,
                                  const TGraphParams &P, float &rho_out, float &sigma_out) {
    int positives = 0;
    float pmax = 0.0f;
    for (int s = 0; s < k; ++s)
        if (dist[s] > 0.0f) {
            ++positives;
            pmax = fmaxf(pmax, dist[s]);
        }
    float rho = 0.0f;
    if (positives > 0) {
        if (P.floor_conn == 0) {
            float first = pmax;
            for (int s = 0; s < k; ++s)
                if (dist[s] > 0.0f && dist[s] < first) first = dist[s];
            rho = first * P.fraction;
        } else {
            const int64_t lo_rank = static_cast<int64_t>(P.floor_conn) - 1, hi_rank = P.floor_conn;
            float before = pmax, after = pmax;
            for (int c = 0; c < k; ++c) {
                const float cand = dist[c];
                if (!(cand > 0.0f)) continue;
                int64_t less = 0, equal = 0;
                for (int o = 0; o < k; ++o) {
                    if (dist[o] > 0.0f && dist[o] < cand) ++less;
                    else if (dist[o] > 0.0f && dist[o] == cand) ++equal;
                }
                if (less <= lo_rank && lo_rank < less + equal) before = cand;
                if (less <= hi_rank && hi_rank < less + equal) after = cand;
            }
            rho = P.interpolate ? before + (after - before) * P.fraction : before;
        }
        if (positives < static_cast<int>(P.min_positive)) rho = pmax;
    }
    float row_sum = 0.0f;
    for (int s = 0; s < k; ++s) row_sum += dist[s];
    const float floor_mean =
        rho > 0.0f ? row_sum / static_cast<float>(static_cast<uint32_t>(k)) : global_mean;
    float lo = 0.0f, hi = 3.4028235e38f, mid = 1.0f;
    for (int it = 0; it < 64; ++it) {
        float sum = 0.0f;
        for (int s = 1; s < k; ++s) {
            const float delta = dist[s] - rho;
            const float safe = mid == 0.0f ? 1.0f : mid;
            const float clamped = delta < 0.0f ? 0.0f : delta;
            float value = expf(-clamped / safe);
            if (delta <= 0.0f) value = 1.0f;
            sum += value;
        }
        if (fabsf(sum - P.target) < 1e-5f) break;
        if (sum > P.target) {
            hi = mid;
            mid = (lo + hi) * 0.5f;
        } else if (hi == 3.4028235e38f) {
            lo = mid;
            mid *= 2.0f;
        } else {
            lo = mid;
            mid = (lo + hi) * 0.5f;
        }
    }
    const float floor = floor_mean * 1e-3f;
    if (mid < floor) mid = floor;
    rho_out = rho;
    sigma_out = mid;
}

__global__ __launch_bounds__(kRowThreads) void tg_rows_kernel(
    const TGraphSegment *segs, int32_t n_segs, int64_t total_rows, const int32_t *idx,
    const float *dist, int64_t k, int64_t n, TGraphParams P, const float *embedding, int64_t d,
    int32_t *tmp_tail, float *tmp_w, int32_t *tmp_cnt) {
    const int64_t row = static_cast<int64_t>(blockIdx.x) * kRowThreads + threadIdx.x;
    if (row >= total_rows) return;
    const TGraphSegment &S = segs[find_segment(segs, n_segs, row)];
    const int kk = static_cast<int>(k);
    const int32_t *ids = idx + row * k;
    const float *dd = dist + row * k;
    float rho, sigma;
    smooth_row(dd, kk, S.state->mean, P, rho, sigma);
    // memberships (fuzzy_graph, graph.rs:252-337 with bipartite = true)
    float mem[kMaxGraphK];
    bool invalid = false;
    for (int s = 0; s < kk; ++s) {
        const int32_t index = ids[s];
        const float delta = dd[s] - rho;
        const float safe = sigma == 0.0f ? 1.0f : sigma;
        const float clamped = delta < 0.0f ? 0.0f : delta;
        float value = expf(-clamped / safe);
        if (dd[s] <= rho || sigma == 0.0f) value = 1.0f;
        if (index < 0 || index >= n) value = 0.0f;
        mem[s] = value;
        if (index < -1 || index >= n) invalid = true;
    }
    // the row's edges: first occurrence of each reference id carries the sum
    // of its duplicates' memberships; kept when positive; sorted by id
    int32_t et[kMaxGraphK];
    float ew[kMaxGraphK];
    int c = 0;
    for (int s = 0; s < kk; ++s) {
        const int32_t tail = ids[s];
        if (!(tail >= 0 && tail < n)) continue;
        bool first = true;
        float forward = 0.0f;
        for (int o = 0; o < kk; ++o)
            if (ids[o] == tail) {
                if (o < s) first = false;
                forward += mem[o];
            }
        if (!first || !(forward > 0.0f)) continue;
        int p = c++;
        while (p > 0 && et[p - 1] > tail) {
            et[p] = et[p - 1];
            ew[p] = ew[p - 1];
            --p;
        }
        et[p] = tail;
        ew[p] = forward;
    }
    uint32_t wmax = 0u;
    for (int j = 0; j < c; ++j) {
        tmp_tail[row * k + j] = et[j];
        tmp_w[row * k + j] = ew[j];
        wmax = max(wmax, __float_as_uint(ew[j]));
    }
    tmp_cnt[row] = c;
    if (wmax) atomicMax(&S.state->wmax_bits, wmax);
    // initial placement (estimator.rs:2753-2798): a copy of the smallest
    // reference with membership exactly 1, else the membership-weighted mean
    int64_t selected = -1;
    bool positive = false;
    for (int s = 0; s < kk; ++s) {
        const int32_t id = ids[s];
        if (id < 0) continue;
        bool seen = false;
        for (int o = 0; o < s; ++o) seen = seen || ids[o] == id;
        if (seen) continue;
        float weight = 0.0f;
        for (int o = 0; o < kk; ++o)
            if (ids[o] == id) weight += mem[o];
        positive = positive || weight > 0.0f;
        if (weight == 1.0f && (selected < 0 || id < selected)) selected = id;
    }
    float *out = S.init + (row - S.row0) * d;
    int32_t errors = invalid ? kTGraphInvalidIds : 0;
    if (!positive) {
        errors |= kTGraphEmptyInitRow;
        for (int64_t cc = 0; cc < d; ++cc) out[cc] = 0.0f;
    } else if (selected >= 0) {
        for (int64_t cc = 0; cc < d; ++cc) out[cc] = embedding[selected * d + cc];
    } else {
        float sum = 0.0f;
        for (int s = 0; s < kk; ++s) sum += mem[s];
        const float div = sum == 0.0f ? 1.0f : sum;
        for (int64_t cc = 0; cc < d; ++cc) {
            float o = 0.0f;
            for (int s = 0; s < kk; ++s) {
                const int64_t id = ids[s];
                if (id < 0 || id >= n) continue;   // zero membership (invalid ids fail the request)
                const float nw = mem[s] / div;
                const float prod = embedding[id * d + cc] * nw;
                o = o + prod;
            }
            out[cc] = o;
        }
    }
    if (errors) atomicOr(&S.state->errors, errors);
}

// Block-wide exclusive scan of one value per thread; total over the block.
__device__ inline uint32_t seg_exclusive_scan(uint32_t v, uint32_t *s_wave, uint32_t &total) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t up = __shfl_up(incl, off, 64);
        if (lane >= off) incl += up;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (int w = 0; w < kSegThreads / 64; ++w) {
        if (w < wave) before += s_wave[w];
        total += s_wave[w];
    }
    __syncthreads();
    return before + incl - v;
}

__global__ __launch_bounds__(kSegThreads) void tg_offsets_kernel(const TGraphSegment *segs,
                                                                const float *tmp_w,
                                                                const int32_t *tmp_cnt,
                                                                int64_t k, float prune_divisor) {
    __shared__ uint32_t s_wave[kSegThreads / 64];
    const TGraphSegment &S = segs[blockIdx.x];
    const int tid = threadIdx.x;
    const float maximum = __uint_as_float(S.state->wmax_bits);
    const float threshold = maximum / prune_divisor;
    uint32_t carry = 0;
    for (int64_t r0 = 0; r0 < S.rows; r0 += kSegThreads) {
        const int64_t r = r0 + tid;
        uint32_t kept = 0;
        if (r < S.rows) {
            const int64_t row = S.row0 + r;
            for (int j = 0; j < tmp_cnt[row]; ++j) {
                const float w = tmp_w[row * k + j];
                if (!(w == 0.0f || w < threshold)) ++kept;
            }
        }
        uint32_t chunk;
        const uint32_t ex = seg_exclusive_scan(kept, s_wave, chunk);
        if (r < S.rows) S.offsets[r] = carry + ex;
        carry += chunk;
    }
    if (tid == 0) {
        S.offsets[S.rows] = carry;
        S.state->edges = carry;
        if (carry == 0) atomicOr(&S.state->errors, kTGraphEmpty);
    }
}

__global__ __launch_bounds__(kRowThreads) void tg_csr_kernel(
    const TGraphSegment *segs, int32_t n_segs, int64_t total_rows, const int32_t *tmp_tail,
    const float *tmp_w, const int32_t *tmp_cnt, int64_t k, float prune_divisor, int32_t epochs,
    float rate) {
    const int64_t row = static_cast<int64_t>(blockIdx.x) * kRowThreads + threadIdx.x;
    if (row >= total_rows) return;
    const TGraphSegment &S = segs[find_segment(segs, n_segs, row)];
    const float wmax = __uint_as_float(S.state->wmax_bits);
    const float threshold = wmax / prune_divisor;
    int64_t p = S.offsets[row - S.row0];
    bool bad = false;
    for (int j = 0; j < tmp_cnt[row]; ++j) {
        const float w = tmp_w[row * k + j];
        if (w == 0.0f || w < threshold) continue;
        const float samples = static_cast<float>(epochs) * (w / wmax);
        const double period = static_cast<double>(epochs) / static_cast<double>(samples);
        const double nperiod = rate == 0.0f ? INFINITY : (period / static_cast<double>(rate)) * 1.5;
        if (!isfinite(period) || period <= 0.0 || (rate > 0.0f && (!isfinite(nperiod) || nperiod <= 0.0)))
            bad = true;
        S.indices[p] = static_cast<uint32_t>(tmp_tail[row * k + j]);
        S.periods[p] = period;
        S.negative_periods[p] = nperiod;
        S.next_positive[p] = period;
        S.next_negative[p] = nperiod;
        ++p;
    }
    if (bad) atomicOr(&S.state->errors, kTGraphBadPeriod);
}

// ------------------------------------------------------------- fit graph
// The fit's fuzzy simplicial set on the device: memberships of every row
// (as above, self excluded), the fuzzy union of each unordered pair emitted
// once (fuzzy_graph's evaluate_edge, graph.rs:252-337), pruning below
// max / divisor, both directions of every kept pair sorted by (head, tail)
// with two stable radix sorts, and the compact CSR with clocks.

__global__ __launch_bounds__(kRowThreads) void fg_rows_kernel(const int32_t *idx, const float *dist,
                                                             int64_t n, int64_t k, TGraphParams P,
                                                             const TGraphState *state,
                                                             float *directed, int32_t *errors) {
    const int64_t row = static_cast<int64_t>(blockIdx.x) * kRowThreads + threadIdx.x;
    if (row >= n) return;
    const int kk = static_cast<int>(k);
    const int32_t *ids = idx + row * k;
    const float *dd = dist + row * k;
    float rho, sigma;
    smooth_row(dd, kk, state->mean, P, rho, sigma);
    bool invalid = false;
    for (int s = 0; s < kk; ++s) {
        const int32_t index = ids[s];
        const float delta = dd[s] - rho;
        const float safe = sigma == 0.0f ? 1.0f : sigma;
        const float clamped = delta < 0.0f ? 0.0f : delta;
        float value = expf(-clamped / safe);
        if (dd[s] <= rho || sigma == 0.0f) value = 1.0f;
        if (index < 0 || index >= n || index == row) value = 0.0f;
        directed[row * k + s] = value;
        if (index < -1 || index >= n) invalid = true;
    }
    if (invalid) atomicOr(errors, kTGraphInvalidIds);
}

// Slot e = row * k + s: the union weight of the pair {row, tail} when this
// slot is the pair's emitter (first occurrence in the row, and the smaller
// id when both rows list each other), else 0.
__global__ __launch_bounds__(kSegThreads) void fg_emit_kernel(const int32_t *idx,
                                                             const float *directed, int64_t n,
                                                             int64_t k, float mix, float *emit_w,
                                                             uint32_t *wmax_bits) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kSegThreads + threadIdx.x;
    if (e >= n * k) return;
    const int64_t row = e / k;
    const int slot = static_cast<int>(e - row * k);
    const int kk = static_cast<int>(k);
    float out = 0.0f;
    const int32_t tail = idx[e];
    if (tail >= 0 && tail < n && tail != row) {
        bool first = true;
        float forward = 0.0f;
        for (int o = 0; o < kk; ++o)
            if (idx[row * k + o] == tail) {
                if (o < slot) first = false;
                forward += directed[row * k + o];
            }
        if (first) {
            float reverse = 0.0f;
            bool has_reverse = false;
            for (int rs = 0; rs < kk; ++rs)
                if (idx[static_cast<int64_t>(tail) * k + rs] == static_cast<int32_t>(row)) {
                    reverse += directed[static_cast<int64_t>(tail) * k + rs];
                    has_reverse = true;
                }
            if (!has_reverse || row <= tail) {
                const float inter = forward * reverse;
                const float uni = (forward + reverse) - inter;
                const float fuzzy = mix * uni + (1.0f - mix) * inter;
                if (fuzzy > 0.0f) out = fuzzy;
            }
        }
    }
    emit_w[e] = out;
    if (out > 0.0f) atomicMax(wmax_bits, __float_as_uint(out));
}

// Kept emissions: 1 where the weight is positive and not below the prune
// threshold.
__global__ __launch_bounds__(kSegThreads) void fg_keep_kernel(const float *emit_w, int64_t m,
                                                             const uint32_t *wmax_bits,
                                                             float divisor, uint32_t *keep) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kSegThreads + threadIdx.x;
    if (e >= m) return;
    const float w = emit_w[e];
    const float threshold = divisor > 0.0f ? __uint_as_float(*wmax_bits) / divisor : 0.0f;
    keep[e] = (w == 0.0f || w < threshold) ? 0u : 1u;
}

// Exclusive scan of m u32 values in three launches (block totals, a scan of
// the totals in one workgroup, block-local scans plus offsets); no
// workgroup waits on another.
constexpr int kScanItems = 4;
constexpr int kScanTile = kSegThreads * kScanItems;

__global__ __launch_bounds__(kSegThreads) void scan_totals_kernel(const uint32_t *in, int64_t m,
                                                                 uint32_t *totals) {
    __shared__ uint32_t s_wave[kSegThreads / 64];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
    uint32_t v = 0;
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t p = base + static_cast<int64_t>(threadIdx.x) * kScanItems + i;
        if (p < m) v += in[p];
    }
    uint32_t total;
    (void)seg_exclusive_scan(v, s_wave, total);
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

__global__ __launch_bounds__(kSegThreads) void scan_block_offsets_kernel(uint32_t *totals,
                                                                        int64_t blocks,
                                                                        uint32_t *grand) {
    __shared__ uint32_t s_wave[kSegThreads / 64];
    uint32_t carry = 0;
    for (int64_t b0 = 0; b0 < blocks; b0 += kSegThreads) {
        const int64_t b = b0 + threadIdx.x;
        const uint32_t v = b < blocks ? totals[b] : 0u;
        uint32_t chunk;
        const uint32_t ex = seg_exclusive_scan(v, s_wave, chunk);
        if (b < blocks) totals[b] = carry + ex;
        carry += chunk;
    }
    if (threadIdx.x == 0) *grand = carry;
}

__global__ __launch_bounds__(kSegThreads) void scan_apply_kernel(const uint32_t *in, int64_t m,
                                                                const uint32_t *offsets,
                                                                uint32_t *out) {
    __shared__ uint32_t s_wave[kSegThreads / 64];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
    uint32_t v[kScanItems], sum = 0;
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t p = base + static_cast<int64_t>(threadIdx.x) * kScanItems + i;
        v[i] = p < m ? in[p] : 0u;
        sum += v[i];
    }
    uint32_t total;
    uint32_t run = offsets[blockIdx.x] + seg_exclusive_scan(sum, s_wave, total);
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t p = base + static_cast<int64_t>(threadIdx.x) * kScanItems + i;
        if (p < m) out[p] = run;
        run += v[i];
    }
}

// Both directions of every kept pair: (min, max) at 2p and (max, min) at 2p + 1.
__global__ __launch_bounds__(kSegThreads) void fg_pairs_kernel(const int32_t *idx,
                                                              const float *emit_w,
                                                              const uint32_t *keep,
                                                              const uint32_t *pos, int64_t m,
                                                              int64_t k, uint32_t *heads,
                                                              uint32_t *tails, float *w) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kSegThreads + threadIdx.x;
    if (e >= m || !keep[e]) return;
    const uint32_t row = static_cast<uint32_t>(e / k);
    const uint32_t tail = static_cast<uint32_t>(idx[e]);
    const uint32_t lo = row < tail ? row : tail, hi = row < tail ? tail : row;
    const int64_t p = 2 * static_cast<int64_t>(pos[e]);
    heads[p] = lo;
    tails[p] = hi;
    heads[p + 1] = hi;
    tails[p + 1] = lo;
    w[p] = emit_w[e];
    w[p + 1] = emit_w[e];
}

__global__ __launch_bounds__(kSegThreads) void gather_u32_kernel(const uint32_t *src,
                                                                const int32_t *perm, int64_t m,
                                                                uint32_t *dst) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kSegThreads + threadIdx.x;
    if (i < m) dst[i] = src[perm[i]];
}

// Row offsets from the sorted heads: offsets[r] = first position with head >= r.
__global__ __launch_bounds__(kSegThreads) void fg_offsets_kernel(const uint32_t *heads_sorted,
                                                                int64_t m, int64_t n,
                                                                uint32_t *offsets) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * kSegThreads + threadIdx.x;
    if (r > n) return;
    int64_t lo = 0, hi = m;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (heads_sorted[mid] < static_cast<uint32_t>(r)) lo = mid + 1;
        else hi = mid;
    }
    offsets[r] = static_cast<uint32_t>(lo);
}

// Column, weight, periods and clocks of edge i of the sorted CSR.
__global__ __launch_bounds__(kSegThreads) void fg_csr_kernel(
    const uint32_t *tails, const float *w, const int32_t *perm, int64_t m,
    const uint32_t *wmax_bits, int32_t epochs, float rate, uint32_t *indices, float *weights,
    double *periods, double *negative_periods, double *next_positive, double *next_negative,
    int32_t *errors) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kSegThreads + threadIdx.x;
    if (i >= m) return;
    const int64_t src = perm[i];
    const float wi = w[src];
    const float wmax = __uint_as_float(*wmax_bits);
    const float samples = static_cast<float>(epochs) * (wi / wmax);
    const double period = static_cast<double>(epochs) / static_cast<double>(samples);
    const double nperiod = rate == 0.0f ? INFINITY : (period / static_cast<double>(rate)) * 1.5;
    if (!isfinite(period) || period <= 0.0 || (rate > 0.0f && (!isfinite(nperiod) || nperiod <= 0.0)))
        atomicOr(errors, kTGraphBadPeriod);
    indices[i] = tails[src];
    weights[i] = wi;
    periods[i] = period;
    negative_periods[i] = nperiod;
    next_positive[i] = period;
    next_negative[i] = nperiod;
}

}  // namespace

TGraphParams transform_graph_params(int64_t k, float local_connectivity) {
    // fuzzy_graph's parameters (graph.rs:21-43) with bandwidth 1
    TGraphParams P{};
    const uint64_t floor_c = static_cast<uint64_t>(local_connectivity);
    P.fraction = local_connectivity - static_cast<float>(floor_c);
    const uint64_t min_pos = floor_c + (local_connectivity > static_cast<float>(floor_c) ? 1 : 0);
    const uint64_t above = static_cast<uint64_t>(k) + 1;
    P.floor_conn = static_cast<uint32_t>(std::min(floor_c, above));
    P.min_positive = static_cast<uint32_t>(std::min(min_pos, above));
    P.interpolate = P.fraction > 1e-5f ? 1 : 0;
    P.target = std::log2(static_cast<float>(k)) * 1.0f;
    return P;
}

void launch_transform_graphs(const TGraphSegment *segs, int32_t n_segs, int64_t total_rows,
                             const int32_t *idx, const float *dist, int64_t k, int64_t n,
                             const TGraphParams &P, const float *embedding, int64_t d,
                             int32_t epochs, float rate, int32_t *tmp_tail, float *tmp_w,
                             int32_t *tmp_cnt, hipStream_t s) {
    if (n_segs == 0 || total_rows == 0) return;
    UMAPHIP_REQUIRE(k >= 1 && k <= kMaxGraphK, UMAPHIP_INVALID_CONFIG,
                    "device transform graphs support 1..%d neighbours", kMaxGraphK);
    const unsigned row_blocks = static_cast<unsigned>((total_rows + kRowThreads - 1) / kRowThreads);
    const float divisor = static_cast<float>(static_cast<double>(epochs));
    hipLaunchKernelGGL(tg_mean_kernel, dim3(static_cast<unsigned>(n_segs)), dim3(kSegThreads), 0, s,
                       segs, dist, k);
    hipLaunchKernelGGL(tg_rows_kernel, dim3(row_blocks), dim3(kRowThreads), 0, s, segs, n_segs,
                       total_rows, idx, dist, k, n, P, embedding, d, tmp_tail, tmp_w, tmp_cnt);
    hipLaunchKernelGGL(tg_offsets_kernel, dim3(static_cast<unsigned>(n_segs)), dim3(kSegThreads), 0,
                       s, segs, tmp_w, tmp_cnt, k, divisor);
    hipLaunchKernelGGL(tg_csr_kernel, dim3(row_blocks), dim3(kRowThreads), 0, s, segs, n_segs,
                       total_rows, tmp_tail, tmp_w, tmp_cnt, k, divisor, epochs, rate);
    HIP_CHECK(hipGetLastError());
}

const char *transform_graph_error(int32_t errors) {
    if (errors & kTGraphInvalidIds) return "fuzzy graph contains missing or invalid neighbor indices";
    if (errors & kTGraphEmpty) return "transform fuzzy graph is empty after positive-edge pruning";
    if (errors & kTGraphEmptyInitRow) return "transform initialization row has no positive memberships";
    if (errors & kTGraphBadPeriod) return "invalid layout period";
    return "";
}

umaphip_status transform_graph_status(int32_t errors) {
    if (errors & kTGraphInvalidIds) return UMAPHIP_NUMERICAL;
    if (errors & kTGraphEmpty) return UMAPHIP_INVALID_INPUT;
    if (errors) return UMAPHIP_NUMERICAL;
    return UMAPHIP_OK;
}

}  // namespace umaphip

using namespace umaphip;

extern "C" {

umaphip_status umaphip_transform_graph(const int32_t *indices, const float *distances, int64_t q,
                                       int64_t k, const float *embedding, int64_t n, int64_t d,
                                       float local_connectivity, int32_t epochs, float rate,
                                       uint32_t *offsets, uint32_t *out_indices, double *periods,
                                       double *negative_periods, double *next_positive,
                                       double *next_negative, float *init, int64_t *n_edges,
                                       void *stream) {
    return guarded([&] {
        UMAPHIP_REQUIRE(indices && distances && embedding && offsets && out_indices && periods &&
                            negative_periods && next_positive && next_negative && init && n_edges,
                        UMAPHIP_INVALID_INPUT, "transform graph arguments are null");
        UMAPHIP_REQUIRE(q > 0 && k > 0 && n > 0 && d > 0, UMAPHIP_INVALID_INPUT,
                        "transform graph shapes must be positive");
        UMAPHIP_REQUIRE(std::isfinite(local_connectivity) && local_connectivity >= 0.0f && epochs > 0 &&
                            std::isfinite(rate) && rate >= 0.0f,
                        UMAPHIP_INVALID_CONFIG, "invalid transform graph settings");
        UMAPHIP_REQUIRE(q * k <= UINT32_MAX, UMAPHIP_OVERFLOW, "CSR sizes");
        hipStream_t s = as_stream(stream);
        DeviceArray<int32_t> tmp_tail(static_cast<size_t>(q * k)), tmp_cnt(static_cast<size_t>(q));
        DeviceArray<float> tmp_w(static_cast<size_t>(q * k));
        DeviceArray<TGraphState> state(1);
        DeviceArray<TGraphSegment> seg(1);
        TGraphSegment S{};
        S.row0 = 0;
        S.rows = q;
        S.offsets = offsets;
        S.indices = out_indices;
        S.periods = periods;
        S.negative_periods = negative_periods;
        S.next_positive = next_positive;
        S.next_negative = next_negative;
        S.init = init;
        S.state = state.ptr;
        HIP_CHECK(hipMemcpyAsync(seg.ptr, &S, sizeof(S), hipMemcpyHostToDevice, s));
        launch_transform_graphs(seg.ptr, 1, q, indices, distances, k, n,
                                transform_graph_params(k, local_connectivity), embedding, d, epochs,
                                rate, tmp_tail.ptr, tmp_w.ptr, tmp_cnt.ptr, s);
        TGraphState host{};
        HIP_CHECK(hipMemcpyAsync(&host, state.ptr, sizeof(host), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        *n_edges = host.edges;
        UMAPHIP_REQUIRE(host.errors == 0, transform_graph_status(host.errors), "%s",
                        transform_graph_error(host.errors));
    });
}

}  // extern "C"

extern "C" {

umaphip_status umaphip_fit_graph(const int32_t *indices, const float *distances, int64_t n,
                                 int64_t k, float local_connectivity, float bandwidth, float mix,
                                 int32_t epochs, float rate, double prune_divisor, uint32_t *offsets,
                                 uint32_t *out_indices, float *weights, double *periods,
                                 double *negative_periods, double *next_positive,
                                 double *next_negative, int64_t *n_edges, void *stream) {
    return guarded([&] {
        UMAPHIP_REQUIRE(indices && distances && offsets && out_indices && weights && periods &&
                            negative_periods && next_positive && next_negative && n_edges,
                        UMAPHIP_INVALID_INPUT, "fit graph arguments are null");
        UMAPHIP_REQUIRE(n > 0 && k > 0 && k <= kMaxGraphK, UMAPHIP_INVALID_INPUT,
                        "fit graph needs rows and 1..%d neighbours", kMaxGraphK);
        UMAPHIP_REQUIRE(std::isfinite(local_connectivity) && local_connectivity >= 0.0f &&
                            std::isfinite(bandwidth) && bandwidth > 0.0f && std::isfinite(mix) &&
                            mix >= 0.0f && mix <= 1.0f && epochs > 0 && std::isfinite(rate) &&
                            rate >= 0.0f,
                        UMAPHIP_INVALID_CONFIG, "invalid fit graph settings");
        UMAPHIP_REQUIRE(2 * n * k <= INT32_MAX, UMAPHIP_OVERFLOW, "fit graph COO indices");
        hipStream_t s = as_stream(stream);
        const int64_t m = n * k;
        TGraphParams P = transform_graph_params(k, local_connectivity);
        P.target = std::log2(static_cast<float>(k)) * bandwidth;
        DeviceArray<TGraphState> state(1);
        DeviceArray<float> directed(m), emit_w(m);
        DeviceArray<uint32_t> keep(m), pos(m);
        const int64_t blocks = (m + kScanTile - 1) / kScanTile;
        DeviceArray<uint32_t> totals(blocks), grand(1);
        const unsigned gm = static_cast<unsigned>((m + kSegThreads - 1) / kSegThreads);
        DeviceArray<double> mean_partial(kMeanBlocks);
        hipLaunchKernelGGL(fg_mean_partial_kernel, dim3(kMeanBlocks), dim3(kSegThreads), 0, s,
                           distances, m, mean_partial.ptr);
        hipLaunchKernelGGL(fg_mean_final_kernel, dim3(1), dim3(64), 0, s, mean_partial.ptr, m,
                           state.ptr);
        hipLaunchKernelGGL(fg_rows_kernel, dim3(static_cast<unsigned>((n + kRowThreads - 1) / kRowThreads)),
                           dim3(kRowThreads), 0, s, indices, distances, n, k, P, state.ptr,
                           directed.ptr, &state.ptr->errors);
        hipLaunchKernelGGL(fg_emit_kernel, dim3(gm), dim3(kSegThreads), 0, s, indices, directed.ptr,
                           n, k, mix, emit_w.ptr, &state.ptr->wmax_bits);
        const float divisor = prune_divisor > 0.0 ? static_cast<float>(prune_divisor) : 0.0f;
        hipLaunchKernelGGL(fg_keep_kernel, dim3(gm), dim3(kSegThreads), 0, s, emit_w.ptr, m,
                           &state.ptr->wmax_bits, divisor, keep.ptr);
        hipLaunchKernelGGL(scan_totals_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kSegThreads),
                           0, s, keep.ptr, m, totals.ptr);
        hipLaunchKernelGGL(scan_block_offsets_kernel, dim3(1), dim3(kSegThreads), 0, s, totals.ptr,
                           blocks, grand.ptr);
        hipLaunchKernelGGL(scan_apply_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kSegThreads), 0,
                           s, keep.ptr, m, totals.ptr, pos.ptr);
        HIP_CHECK(hipGetLastError());
        uint32_t kept = 0;
        TGraphState hs{};
        HIP_CHECK(hipMemcpyAsync(&kept, grand.ptr, sizeof(kept), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipMemcpyAsync(&hs, state.ptr, sizeof(hs), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        UMAPHIP_REQUIRE(!(hs.errors & kTGraphInvalidIds), UMAPHIP_NUMERICAL,
                        "fuzzy graph contains missing or invalid neighbor indices");
        UMAPHIP_REQUIRE(kept > 0, UMAPHIP_INVALID_INPUT,
                        "fuzzy graph is empty after positive-edge pruning");
        const int64_t E = 2 * static_cast<int64_t>(kept);
        DeviceArray<uint32_t> heads(E), tails(E), keys_sorted(E), keys2(E);
        DeviceArray<float> w(E);
        DeviceArray<int32_t> perm1(E), perm(E);
        DeviceArray<uint8_t> ws(sort_pairs_workspace(E));
        hipLaunchKernelGGL(fg_pairs_kernel, dim3(gm), dim3(kSegThreads), 0, s, indices, emit_w.ptr,
                           keep.ptr, pos.ptr, m, k, heads.ptr, tails.ptr, w.ptr);
        // (head, tail) order: stable by tail, then stable by head
        launch_sort_pairs(ws.ptr, ws.count, tails.ptr, keys_sorted.ptr, nullptr, perm1.ptr, E, s);
        const unsigned ge = static_cast<unsigned>((E + kSegThreads - 1) / kSegThreads);
        hipLaunchKernelGGL(gather_u32_kernel, dim3(ge), dim3(kSegThreads), 0, s, heads.ptr, perm1.ptr,
                           E, keys2.ptr);
        launch_sort_pairs(ws.ptr, ws.count, keys2.ptr, keys_sorted.ptr, perm1.ptr, perm.ptr, E, s);
        hipLaunchKernelGGL(fg_offsets_kernel, dim3(static_cast<unsigned>((n + 1 + kSegThreads - 1) / kSegThreads)),
                           dim3(kSegThreads), 0, s, keys_sorted.ptr, E, n, offsets);
        hipLaunchKernelGGL(fg_csr_kernel, dim3(ge), dim3(kSegThreads), 0, s, tails.ptr, w.ptr,
                           perm.ptr, E, &state.ptr->wmax_bits, epochs, rate, out_indices, weights,
                           periods, negative_periods, next_positive, next_negative,
                           &state.ptr->errors);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(&hs, state.ptr, sizeof(hs), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        UMAPHIP_REQUIRE(!(hs.errors & kTGraphBadPeriod), UMAPHIP_NUMERICAL, "invalid layout period");
        *n_edges = E;
    });
}

}  // extern "C"
