/*
 * CPU restatement of cuscan's modern UMAP layout path -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker / baseline. The shipped package
 * (cuscan_amd) never links or calls it.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off). Floating
 * point contraction is disabled because the reference's Rust host code does
 * not contract a*b+c into fused multiply-adds, and the layout clocks
 * (next_negative += count * period) must be bit-exact.
 *
 * Parity status (how each group is pinned; see tests/test_oracle_*.py):
 *  - schedules, compact graph, clocks, forces, gradient kernel, lazy Adam,
 *    momentum blocks: pinned by the reference's frozen umap-learn fixture
 *    (crates/umap-core/tests/fixtures/modern_layout_upstream.json, compacted
 *    into tests/golden/) and by the inline closed-form values of
 *    crates/umap-core/src/modern_layout.rs:1243-1920.
 *  - negative indices / windows / retries: pinned by closed forms (u128
 *    modular identities, window endpoints) and brute force on tiny inputs.
 *  - projection sort / ranks: pinned by numpy's stable argsort (a textbook
 *    routine) on the radix key, and by searchsorted for lower bounds.
 *  - adaptation ratio / scale: pinned by closed forms (constant log ratio,
 *    numpy linear quantiles) and the reference's inline ratio test values.
 *  - exact kNN: pinned by brute force (numpy) with index tie-breaks.
 *  - fuzzy graph: pinned by an independent numpy restatement on small inputs
 *    and closed-form memberships.
 *  - fit_curve: pinned by umap-learn's published (a, b) for the defaults.
 *  - ChaCha12 block: pinned by the RFC 7539 ChaCha20 block test vector
 *    (same quarter-round, 20 rounds) -- the round function is pinned.
 *    The StdRng stream conventions (seed_from_u64's PCG32 expansion, 4-block
 *    buffering, u64 assembly), SliceRandom::shuffle's IncreasingUniform index
 *    draws, Canon's range sampling and the float conversions restate
 *    rand 0.10.3 / rand_core 0.10.1 / chacha20 0.10.2 from their published
 *    algorithms; those crates are not present here, so the RNG STREAM is
 *    PARITY UNPINNED (only self-consistency and distributional checks).
 */
#include "umap_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ==================================================================== RNG */
/* ChaCha quarter round / block function (Bernstein 2008; RFC 7539 s2.1-2.3). */
static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

#define QR(a, b, c, d)                 \
    a += b; d ^= a; d = rotl32(d, 16); \
    c += d; b ^= c; b = rotl32(b, 12); \
    a += b; d ^= a; d = rotl32(d, 8);  \
    c += d; b ^= c; b = rotl32(b, 7);

void orc_chacha_block(const uint32_t input[16], int rounds, uint32_t out[16]) {
    uint32_t x[16];
    memcpy(x, input, sizeof(x));
    for (int r = 0; r < rounds; r += 2) {
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + input[i];
}

/* rand_core SeedableRng::seed_from_u64: PCG32 (XSH RR) fills the 32-byte
 * seed four bytes at a time, little endian. */
static uint32_t pcg32_step(uint64_t *state) {
    *state = *state * 6364136223846793005ULL + 11634580027462260723ULL;
    uint64_t s = *state;
    uint32_t xorshifted = (uint32_t)(((s >> 18) ^ s) >> 27);
    uint32_t rot = (uint32_t)(s >> 59);
    return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
}

void orc_rng_seed(orc_rng *rng, uint64_t seed) {
    uint64_t state = seed;
    for (int i = 0; i < 8; ++i) rng->key[i] = pcg32_step(&state);
    rng->block = 0;
    rng->idx = 64;
}

static void rng_refill(orc_rng *rng) {
    uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    for (int i = 0; i < 8; ++i) in[4 + i] = rng->key[i];
    for (int b = 0; b < 4; ++b) {
        uint64_t ctr = rng->block + (uint64_t)b;
        in[12] = (uint32_t)ctr;
        in[13] = (uint32_t)(ctr >> 32);
        in[14] = 0;
        in[15] = 0;
        orc_chacha_block(in, 12, rng->buf + 16 * b);
    }
    rng->block += 4;
    rng->idx = 0;
}

uint32_t orc_rng_u32(orc_rng *rng) {
    if (rng->idx >= 64) rng_refill(rng);
    return rng->buf[rng->idx++];
}

uint64_t orc_rng_u64(orc_rng *rng) {
    uint64_t lo = orc_rng_u32(rng);
    uint64_t hi = orc_rng_u32(rng);
    return (hi << 32) | lo;
}

/* StandardUniform f64: 53 high bits times 2^-53. */
double orc_rng_f64(orc_rng *rng) {
    uint64_t v = orc_rng_u64(rng) >> 11;
    return (double)v * (1.0 / 9007199254740992.0);
}

/* UniformInt<u32>::sample_single over [0, bound): Canon's widening multiply
 * with one extra draw when the low half may carry. */
uint32_t orc_rng_range_u32(orc_rng *rng, uint32_t bound) {
    if (bound == 0) return orc_rng_u32(rng); /* full range (unused) */
    uint64_t prod = (uint64_t)orc_rng_u32(rng) * (uint64_t)bound;
    uint32_t result = (uint32_t)(prod >> 32);
    uint32_t lo = (uint32_t)prod;
    if (lo > (uint32_t)(0u - bound)) {
        uint64_t prod2 = (uint64_t)orc_rng_u32(rng) * (uint64_t)bound;
        uint32_t hi2 = (uint32_t)(prod2 >> 32);
        if ((uint64_t)lo + (uint64_t)hi2 > 0xFFFFFFFFull) result += 1;
    }
    return result;
}

/* UniformFloat<f32>::sample_single: 23 random mantissa bits in [1,2),
 * minus one, times the scale, plus low; retried (with a one-ulp smaller
 * scale) in the measure-zero case that rounding reaches high. */
float orc_rng_range_f32(orc_rng *rng, float low, float high) {
    float scale = high - low;
    for (;;) {
        uint32_t bits = (orc_rng_u32(rng) >> 9) | 0x3F800000u;
        float v12;
        memcpy(&v12, &bits, 4);
        float v01 = v12 - 1.0f;
        float prod = v01 * scale;
        float res = prod + low;
        if (res < high) return res;
        uint32_t sb;
        memcpy(&sb, &scale, 4);
        sb -= 1;
        memcpy(&scale, &sb, 4);
    }
}

/* rand::seq::IncreasingUniform bound helper: the largest product
 * m*(m+1)*...*(m+c-1) that fits u32, and c. */
static void bound_u32(uint32_t m, uint32_t *bound, uint32_t *count) {
    uint64_t product = m;
    uint64_t current = (uint64_t)m + 1;
    for (;;) {
        uint64_t p = product * current;
        if (p > 0xFFFFFFFFull) {
            *bound = (uint32_t)product;
            *count = (uint32_t)(current - m);
            return;
        }
        product = p;
        current += 1;
    }
}

/* SliceRandom::shuffle == partial_shuffle(len): Durstenfeld order with
 * swap(i, index) for i = 0..len and index drawn by IncreasingUniform(n=0). */
void orc_rng_shuffle_i32(orc_rng *rng, int32_t *a, size_t n) {
    if (n <= 1) return;
    uint32_t cn = 0;           /* IncreasingUniform.n */
    uint32_t chunk = 0;
    uint32_t remaining = 1;    /* n == 0 => first index is 0 without a draw */
    for (size_t i = 0; i < n; ++i) {
        uint32_t next_n = cn + 1;
        uint32_t next_remaining;
        if (remaining == 0) {
            uint32_t bound, count;
            bound_u32(next_n, &bound, &count);
            chunk = orc_rng_range_u32(rng, bound);
            next_remaining = count - 1;
        } else {
            next_remaining = remaining - 1;
        }
        uint32_t index;
        if (next_remaining == 0) {
            index = chunk;
        } else {
            index = chunk % next_n;
            chunk /= next_n;
        }
        remaining = next_remaining;
        cn = next_n;
        int32_t t = a[i];
        a[i] = a[index];
        a[index] = t;
    }
}

/* recursive_init::gaussian_pair (umap-core/src/recursive_init/mod.rs:554-558). */
void orc_gaussian_pair(orc_rng *rng, double *first, double *second) {
    double u1 = orc_rng_f64(rng);
    double radius = sqrt(-2.0 * log(1.0 - u1));
    double angle = 6.283185307179586 * orc_rng_f64(rng);
    *first = radius * cos(angle);
    *second = radius * sin(angle);
}

/* ====================================================== host building blocks */
/* adam_schedule (umap-core/src/modern_layout.rs:110-134). */
void orc_adam_schedule(int epochs, double initial_alpha, int good, double *alpha,
                       double *momentum) {
    int warmup = good ? (epochs / 8 < 100 ? epochs / 8 : 100)
                      : (epochs / 2 < 100 ? epochs / 2 : 100);
    for (int n = 0; n < epochs; ++n) {
        double a;
        if (n < warmup) {
            double k = good ? 0.5 : 2.0;
            double r = 1.0 - (double)n / (double)warmup;
            a = (k * initial_alpha - 0.1) * (r * r) + 0.1;
        } else {
            a = 0.15 * (1.0 - (double)(n - warmup) / (double)(epochs - warmup)) + 0.05;
        }
        alpha[n] = a;
        momentum[n] = 0.0;
    }
}

/* momentum_schedule (umap-core/src/modern_layout.rs:137-165). */
void orc_momentum_schedule(int epochs, double initial_alpha, int good, double *alpha,
                           double *momentum) {
    int w = epochs / 8 > 200 ? epochs / 8 : 200;
    int warmup = epochs < w ? epochs : w;
    for (int n = 0; n < epochs; ++n) {
        double r;
        if (good) {
            if (n < warmup) {
                double t = 1.0 - (double)n / (double)warmup;
                r = t * t;
            } else {
                r = 0.0;
            }
        } else {
            double t = 1.0 - (double)n / (double)epochs;
            r = 0.25 * (t * t);
        }
        alpha[n] = good ? 0.25 * initial_alpha * r + 0.005 : initial_alpha * r;
        momentum[n] = 0.5 * (1.0 - r);
    }
}

/* negative_index (umap-core/src/modern_layout.rs:433-447), u128 arithmetic. */
uint64_t orc_negative_index(uint64_t edge, uint64_t epoch, uint64_t sample, uint64_t m) {
    unsigned __int128 mm = m;
    unsigned __int128 factor = ((epoch % mm) + (sample % mm) + 1) % mm;
    return (uint64_t)(((edge % mm) * factor) % mm);
}

/* effective_negative_range (:450-457). */
uint64_t orc_effective_negative_range(uint64_t raw, uint64_t refs) {
    uint64_t r = raw > 200 ? raw : 200;
    return refs < r ? refs : r;
}

/* window_negative_index (:461-478). */
uint64_t orc_window_negative_index(uint64_t edge, uint64_t epoch, uint64_t sample,
                                   uint64_t rank, uint64_t refs, uint64_t range) {
    uint64_t half = range / 2;
    uint64_t start = rank > half ? rank - half : 0;
    if (start > refs - range) start = refs - range;
    return start + orc_negative_index(edge, epoch, sample, range);
}

/* adaptation_due (:518-534). */
int orc_adaptation_due(int fit, int adam, uint64_t samples, uint64_t raw_range,
                       uint64_t refs, uint64_t epoch, uint64_t total) {
    return fit && adam && samples > 0 && raw_range < refs && epoch >= 10 &&
           epoch <= total / 2 && epoch % 10 == 0;
}

/* resolved_negative_scale (:537-561). */
float orc_resolved_negative_scale(uint64_t raw_range, uint64_t refs, int has_explicit,
                                  float explicit_scale, int adaptation_enabled) {
    if (has_explicit) return explicit_scale;
    if (adaptation_enabled) {
        uint64_t q = raw_range < refs ? raw_range : refs;
        return (float)pow((double)q / (double)refs, 0.75);
    }
    return (float)orc_effective_negative_range(raw_range, refs) / (float)refs;
}

/* edge_periods (:188-208). */
static int edge_periods(float weight, float max_weight, int epochs, float rate,
                        double *period, double *negative_period) {
    float samples = (float)epochs * (weight / max_weight);
    double p = (double)epochs / (double)samples;
    double np = rate == 0.0f ? INFINITY : (p / (double)rate) * 1.5;
    if (!isfinite(p) || p <= 0.0 || (rate > 0.0f && (!isfinite(np) || np <= 0.0))) return -1;
    *period = p;
    *negative_period = np;
    return 0;
}

typedef struct {
    int64_t row, col, pos;
    float w;
} coo_entry;

static int coo_cmp(const void *x, const void *y) {
    const coo_entry *a = (const coo_entry *)x, *b = (const coo_entry *)y;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    if (a->col != b->col) return a->col < b->col ? -1 : 1;
    if (a->pos != b->pos) return a->pos < b->pos ? -1 : 1;
    return 0;
}

/* compact_graph (umap-core/src/modern_layout.rs:303-430). Duplicates are
 * summed in input order (the reference's unstable sort leaves their order
 * unspecified; fuzzy-graph inputs never contain duplicates). */
long orc_compact_graph(int64_t rows, int64_t columns, int64_t n_edges, const int64_t *row,
                       const int64_t *column, const float *weight, int epochs, float rate,
                       uint32_t *offsets, uint32_t *indices, float *weights, double *periods,
                       double *negative_periods, double *next_positive,
                       double *next_negative) {
    if (!isfinite(rate) || rate < 0.0f || epochs == 0) return -2; /* InvalidConfig */
    coo_entry *s = (coo_entry *)malloc(sizeof(coo_entry) * (size_t)(n_edges > 0 ? n_edges : 1));
    if (!s) return -4;
    int64_t count = 0;
    for (int64_t i = 0; i < n_edges; ++i) {
        if (!isfinite(weight[i]) || weight[i] < 0.0f) { free(s); return -1; }
        if (weight[i] == 0.0f) continue;
        if (row[i] < 0 || row[i] >= rows || column[i] < 0 || column[i] >= columns) {
            free(s);
            return -1;
        }
        s[count].row = row[i];
        s[count].col = column[i];
        s[count].pos = i;
        s[count].w = weight[i];
        ++count;
    }
    qsort(s, (size_t)count, sizeof(coo_entry), coo_cmp);
    int64_t merged = 0;
    for (int64_t i = 0; i < count; ++i) {
        if (merged > 0 && s[merged - 1].row == s[i].row && s[merged - 1].col == s[i].col) {
            s[merged - 1].w += s[i].w;
            if (!isfinite(s[merged - 1].w)) { free(s); return -3; }
            continue;
        }
        s[merged++] = s[i];
    }
    if (merged == 0) { free(s); return -1; }
    float max_w = 0.0f;
    for (int64_t i = 0; i < merged; ++i)
        if (s[i].w > max_w) max_w = s[i].w;
    for (int64_t r = 0; r <= rows; ++r) offsets[r] = 0;
    for (int64_t i = 0; i < merged; ++i) {
        double p, np;
        if (edge_periods(s[i].w, max_w, epochs, rate, &p, &np)) { free(s); return -3; }
        offsets[s[i].row + 1] += 1;
        indices[i] = (uint32_t)s[i].col;
        weights[i] = s[i].w;
        periods[i] = p;
        negative_periods[i] = np;
        next_positive[i] = p;
        next_negative[i] = np;
    }
    for (int64_t r = 0; r < rows; ++r) offsets[r + 1] += offsets[r];
    free(s);
    return (long)merged;
}

/* advance_clock (umap-core/src/modern_layout.rs:662-691). */
long orc_advance_clock(double *next_positive, double *next_negative, double period,
                       double negative_period, uint64_t epoch) {
    double now = (double)epoch;
    if (*next_positive > now) return -1;
    if (negative_period == INFINITY) {
        *next_positive = *next_positive + period;
        return 0;
    }
    double k = trunc((now - *next_negative) / negative_period);
    double np = *next_positive + period;
    double product = k * negative_period;
    double nn = *next_negative + product;
    *next_positive = np;
    *next_negative = nn;
    return k > 0.0 ? (long)k : 0;
}

/* accumulate_force (umap-core/src/modern_layout.rs:696-742). */
void orc_accumulate_force(float *gradient, const float *source, const float *target, int d,
                          float a, float b, float gamma, float scale, int small,
                          int negative) {
    float q = 0.0f;
    for (int c = 0; c < d; ++c) {
        float z = source[c] - target[c];
        float zz = z * z;
        q += zz;
    }
    if (q <= 0.0f) return;
    float coefficient;
    if (negative) {
        if (gamma <= 0.0f) return;
        float s = small ? 1.0f : scale;
        float num = s * 2.0f;
        num = num * gamma;
        num = num * b;
        float pb = powf(q, b);
        float den1 = 0.001f + q;
        float apb = a * pb;
        float den2 = 1.0f + apb;
        float den = den1 * den2;
        float c = num / den;
        if (c <= 0.0f) return;
        float cc = c * c;
        float norm = sqrtf(cc * q);
        if (norm == 0.0f) return;
        float g16 = gamma < 16.0f ? gamma : 16.0f;
        float t = c * g16;
        t = t * tanhf(norm / gamma);
        coefficient = t / norm;
    } else {
        float num = -2.0f * a;
        num = num * b;
        num = num * powf(q, b - 1.0f);
        float apb = a * powf(q, b);
        float c = num / (1.0f + apb);
        coefficient = c * (small ? 0.5f : 1.0f);
    }
    for (int c = 0; c < d; ++c) {
        float diff = source[c] - target[c];
        float add = coefficient * diff;
        gradient[c] += add;
    }
}

/* ================================================================= epoch */
/* Coordinates follow the kernel's strided access
 * (kernels/standard_umap/modern_layout.rs:17-58). */
static inline float head_at(const orc_epoch_args *A, int64_t row, int64_t c) {
    return A->head[row * A->head_s0 + c * A->head_s1];
}
static inline float target_at(const orc_epoch_args *A, int64_t row, int64_t c) {
    return A->shared ? A->head[row * A->head_s0 + c * A->head_s1]
                     : A->tail[row * A->tail_s0 + c * A->tail_s1];
}

/* excluded (kernels/standard_umap/modern_layout.rs:63-89). */
static int excluded(const orc_epoch_args *A, int64_t begin, int64_t stop, int64_t i,
                    int64_t j) {
    int64_t lo = begin, hi = stop;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if ((int64_t)A->indices[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    int hit = 0;
    if (lo < stop) hit = (int64_t)A->indices[lo] == j;
    if (A->shared) hit = hit || i == j;
    return hit;
}

/* Negative candidate iterator: restates the three selection branches of
 * gradient_kernel (kernels/standard_umap/modern_layout.rs:177-250). */
typedef struct {
    const orc_epoch_args *A;
    int64_t i, e, begin, stop;
    uint64_t n64, range, count, first;
    int small, mode; /* mode 0: small retry, 1: small plain, 2: large window */
    uint64_t attempt, attempts, accepted, base, p;
} neg_iter;

static void neg_init(neg_iter *it, const orc_epoch_args *A, int64_t i, int64_t e,
                     int64_t begin, int64_t stop, int32_t count, int small) {
    it->A = A;
    it->i = i;
    it->e = e;
    it->begin = begin;
    it->stop = stop;
    it->n64 = (uint64_t)A->n;
    it->range = A->negative_range;
    it->count = (uint64_t)count;
    it->small = small;
    it->p = 0;
    it->first = 0;
    it->attempt = 0;
    it->accepted = 0;
    if (small) {
        if (A->exclude) {
            it->mode = 0;
            it->attempts = it->count * it->n64;
            it->base = ((uint64_t)e * ((uint64_t)A->epoch + 1u)) % it->n64;
        } else {
            it->mode = 1;
        }
    } else {
        it->mode = 2;
        it->first = 0;
        if (it->range < it->n64) {
            uint64_t rank = (uint64_t)(uint32_t)A->source_ranks[i];
            uint64_t half = it->range / 2;
            if (rank > half) it->first = rank - half;
            if (it->first > it->n64 - it->range) it->first = it->n64 - it->range;
        }
    }
}

/* Returns 1 with the next candidate (and whether it is kept), 0 when done. */
static int neg_next(neg_iter *it, int64_t *j, int *kept) {
    const orc_epoch_args *A = it->A;
    if (it->mode == 0) {
        if (!(it->accepted < it->count && it->attempt < it->attempts)) return 0;
        uint64_t rank = (it->base + it->attempt % it->n64) % it->n64;
        it->attempt += 1;
        *j = (int64_t)(uint32_t)A->negative_order[rank];
        *kept = !excluded(A, it->begin, it->stop, it->i, *j);
        if (*kept) it->accepted += 1;
        return 1;
    }
    if (it->p >= it->count) return 0;
    uint64_t factor = (uint64_t)A->epoch + it->p + 1u;
    it->p += 1;
    if (it->mode == 1) {
        uint64_t hash = ((uint64_t)it->e * factor) % it->n64;
        *j = (int64_t)(uint32_t)A->negative_order[hash];
        *kept = 1;
        return 1;
    }
    uint64_t selection = ((uint64_t)it->e * factor) % it->range;
    *j = (int64_t)(uint32_t)A->negative_order[it->first + selection];
    *kept = A->exclude ? !excluded(A, it->begin, it->stop, it->i, *j) : 1;
    return 1;
}

/* accumulate (kernels/standard_umap/modern_layout.rs:38-58). */
static void accumulate(const orc_epoch_args *A, int64_t i, int64_t j, float coefficient) {
    for (int64_t c = 0; c < A->d; ++c) {
        float y = target_at(A, j, c);
        float diff = head_at(A, i, c) - y;
        float add = coefficient * diff;
        float *u = &A->updates[i * A->upd_s0 + c * A->upd_s1];
        *u = *u + add;
    }
}

static float distance(const orc_epoch_args *A, int64_t i, int64_t j) {
    float q = 0.0f;
    for (int64_t c = 0; c < A->d; ++c) {
        float y = target_at(A, j, c);
        float z = head_at(A, i, c) - y;
        float zz = z * z;
        q += zz;
    }
    return q;
}

/* repel (kernels/standard_umap/modern_layout.rs:91-126). */
static void repel(const orc_epoch_args *A, int64_t i, int64_t j, float s, float multiplier) {
    float q = distance(A, i, j);
    float gamma = A->gamma, a = A->a, b = A->b;
    if (q > 0.0f && gamma > 0.0f) {
        float num = s * 2.0f;
        num = num * gamma;
        num = num * b;
        float den1 = 0.001f + q;
        float apb = a * powf(q, b);
        float den2 = 1.0f + apb;
        float den = den1 * den2;
        float coeff = num / den;
        if (coeff > 0.0f) {
            float cc = coeff * coeff;
            float norm = sqrtf(cc * q);
            if (norm > 0.0f) {
                float g16 = gamma < 16.0f ? gamma : 16.0f;
                float clip = g16 * tanhf(norm / gamma);
                clip = clip / norm;
                float mc = multiplier * coeff;
                accumulate(A, i, j, mc * clip);
            }
        }
    }
}

/* gradient_kernel (kernels/standard_umap/modern_layout.rs:128-257) for the
 * slots [start, end) of source_order. Rows are distinct within a launch, so
 * slots run in parallel without races. */
void orc_gradient_block(const orc_epoch_args *A, int64_t start, int64_t end, float multiplier,
                        int threads) {
    int small = A->h <= 1024;
    (void)threads;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads > 0 ? threads : 1)
    for (int64_t slot = start; slot < end; ++slot) {
        int64_t i = (int64_t)(uint32_t)A->source_order[slot];
        int64_t begin = A->offsets[i], stop = A->offsets[i + 1];
        for (int64_t e = begin; e < stop; ++e) {
            if (!(A->next_positive[e] <= (double)A->epoch)) continue;
            int64_t j = A->indices[e];
            float q = distance(A, i, j);
            if (q > 0.0f) {
                float factor = small ? 0.5f : 1.0f;
                float fm = factor * multiplier;
                float num = -2.0f * A->a;
                num = num * A->b;
                num = num * powf(q, A->b - 1.0f);
                float top = fm * num;
                float apb = A->a * powf(q, A->b);
                float coeff = top / (1.0f + apb);
                accumulate(A, i, j, coeff);
            }
            A->next_positive[e] = A->next_positive[e] + A->periods[e];
            if (A->negatives) {
                double t = ((double)A->epoch - A->next_negative[e]) / A->negative_periods[e];
                int32_t count = (int32_t)t;
                if (count > 0) {
                    neg_iter it;
                    neg_init(&it, A, i, e, begin, stop, count, small);
                    int64_t jj;
                    int kept;
                    while (neg_next(&it, &jj, &kept))
                        if (kept) repel(A, i, jj, small ? 1.0f : A->scale, multiplier);
                }
                double product = (double)count * A->negative_periods[e];
                A->next_negative[e] = A->next_negative[e] + product;
            }
        }
    }
}

/* The same gradient computed in float64 from the same f32 inputs. */
/* Euclidean norm of one term coeff * (head_i - target_j), d components. */
static double term_norm64(const orc_epoch_args *A, int64_t i, int64_t j, double coeff) {
    double s = 0.0;
    for (int64_t c = 0; c < A->d; ++c) {
        double t = coeff * ((double)head_at(A, i, c) - (double)target_at(A, j, c));
        s += t * t;
    }
    return sqrt(s);
}

static void gradient64(const orc_epoch_args *A, int64_t start, int64_t end, double multiplier,
                       double *g, double *mag) {
    int small = A->h <= 1024;
    double a = A->a, b = A->b, gamma = A->gamma;
    int64_t d = A->d;
    for (int64_t slot = start; slot < end; ++slot) {
        int64_t i = (int64_t)(uint32_t)A->source_order[slot];
        int64_t begin = A->offsets[i], stop = A->offsets[i + 1];
        for (int64_t e = begin; e < stop; ++e) {
            if (!(A->next_positive[e] <= (double)A->epoch)) continue;
            int64_t j = A->indices[e];
            double q = 0.0;
            for (int64_t c = 0; c < d; ++c) {
                double z = (double)head_at(A, i, c) - (double)target_at(A, j, c);
                q += z * z;
            }
            if (q > 0.0) {
                double coeff = (small ? 0.5 : 1.0) * multiplier * (-2.0 * a * b * pow(q, b - 1.0)) /
                               (1.0 + a * pow(q, b));
                for (int64_t c = 0; c < d; ++c)
                    g[i * d + c] += coeff * ((double)head_at(A, i, c) - (double)target_at(A, j, c));
                if (mag) mag[i] += term_norm64(A, i, j, coeff);
            }
            A->next_positive[e] = A->next_positive[e] + A->periods[e];
            if (A->negatives) {
                double t = ((double)A->epoch - A->next_negative[e]) / A->negative_periods[e];
                int32_t count = (int32_t)t;
                if (count > 0) {
                    neg_iter it;
                    neg_init(&it, A, i, e, begin, stop, count, small);
                    int64_t jj;
                    int kept;
                    double s = small ? 1.0 : (double)A->scale;
                    while (neg_next(&it, &jj, &kept)) {
                        if (!kept) continue;
                        double qq = 0.0;
                        for (int64_t c = 0; c < d; ++c) {
                            double z = (double)head_at(A, i, c) - (double)target_at(A, jj, c);
                            qq += z * z;
                        }
                        if (!(qq > 0.0 && gamma > 0.0)) continue;
                        double coeff = s * 2.0 * gamma * b / ((0.001 + qq) * (1.0 + a * pow(qq, b)));
                        if (!(coeff > 0.0)) continue;
                        double norm = sqrt(coeff * coeff * qq);
                        if (!(norm > 0.0)) continue;
                        double clip = fmin(gamma, 16.0) * tanh(norm / gamma) / norm;
                        double f = multiplier * coeff * clip;
                        for (int64_t c = 0; c < d; ++c)
                            g[i * d + c] +=
                                f * ((double)head_at(A, i, c) - (double)target_at(A, jj, c));
                        if (mag) mag[i] += term_norm64(A, i, jj, f);
                    }
                }
                double product = (double)count * A->negative_periods[e];
                A->next_negative[e] = A->next_negative[e] + product;
            }
        }
    }
}

void orc_gradient_block64(const orc_epoch_args *A, int64_t start, int64_t end,
                          double multiplier, double *g) {
    gradient64(A, start, end, multiplier, g, NULL);
}

void orc_gradient_block64_mag(const orc_epoch_args *A, int64_t start, int64_t end,
                              double multiplier, double *g, double *magnitude) {
    gradient64(A, start, end, multiplier, g, magnitude);
}

/* adam_kernel (kernels/standard_umap/modern_layout.rs:495-523). */
void orc_adam_apply(const orc_epoch_args *A, int threads) {
    int64_t total = A->h * A->d;
    (void)threads;
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
    for (int64_t idx = 0; idx < total; ++idx) {
        int64_t row = idx / A->d, col = idx % A->d;
        float *u = &A->updates[row * A->upd_s0 + col * A->upd_s1];
        float *h = &A->head[row * A->head_s0 + col * A->head_s1];
        float g = *u;
        if (g != 0.0f) {
            float one_b1 = 1.0f - 0.9f;
            float one_b2 = 1.0f - 0.99f;
            float first = 0.9f * A->m[idx];
            float t1 = one_b1 * g;
            first = first + t1;
            float second = 0.99f * A->v[idx];
            float t2 = one_b2 * g;
            t2 = t2 * g;
            second = second + t2;
            A->m[idx] = first;
            A->v[idx] = second;
            float mh = first / A->correction1;
            float num = A->alpha * mh;
            float den = sqrtf(second / A->correction2) + 0.0001f;
            float step = num / den;
            *h = *h + step;
        }
        *u = 0.0f;
    }
}

/* apply_kernel (:525-543). */
void orc_momentum_apply_block(const orc_epoch_args *A, int64_t start, int64_t end) {
    for (int64_t slot = start; slot < end; ++slot) {
        int64_t row = (int64_t)(uint32_t)A->source_order[slot];
        for (int64_t c = 0; c < A->d; ++c) {
            float *h = &A->head[row * A->head_s0 + c * A->head_s1];
            *h = *h + A->updates[row * A->upd_s0 + c * A->upd_s1];
        }
    }
}

/* decay_kernel (:545-554). */
void orc_momentum_decay(const orc_epoch_args *A) {
    for (int64_t row = 0; row < A->h; ++row)
        for (int64_t c = 0; c < A->d; ++c) {
            float *u = &A->updates[row * A->upd_s0 + c * A->upd_s1];
            *u = *u * A->mu;
        }
}

/* modern_epoch (kernels/standard_umap/modern_layout.rs:925-1026). */
void orc_modern_epoch(const orc_epoch_args *A, int threads) {
    if (A->momentum) {
        int64_t rows = A->h;
        int64_t block = rows <= 1024 ? rows : (rows / 2 > 1 ? rows / 2 : 1);
        for (int64_t start = 0; start < rows;) {
            int64_t end = start + block < rows ? start + block : rows;
            orc_gradient_block(A, start, end, A->alpha, threads);
            orc_momentum_apply_block(A, start, end);
            start = end;
        }
        orc_momentum_decay(A);
    } else {
        orc_gradient_block(A, 0, A->h, 1.0f, threads);
        orc_adam_apply(A, threads);
    }
}

int64_t orc_trace_negatives(const orc_epoch_args *A, int64_t capacity, int64_t *edge_out,
                            int32_t *cand_out, uint8_t *kept_out, int64_t *fire_count) {
    int small = A->h <= 1024;
    int64_t written = 0, fires = 0;
    for (int64_t slot = 0; slot < A->h; ++slot) {
        int64_t i = (int64_t)(uint32_t)A->source_order[slot];
        int64_t begin = A->offsets[i], stop = A->offsets[i + 1];
        for (int64_t e = begin; e < stop; ++e) {
            if (!(A->next_positive[e] <= (double)A->epoch)) continue;
            ++fires;
            if (!A->negatives) continue;
            double t = ((double)A->epoch - A->next_negative[e]) / A->negative_periods[e];
            int32_t count = (int32_t)t;
            if (count <= 0) continue;
            neg_iter it;
            neg_init(&it, A, i, e, begin, stop, count, small);
            int64_t jj;
            int kept;
            while (neg_next(&it, &jj, &kept)) {
                if (written < capacity) {
                    edge_out[written] = e;
                    cand_out[written] = (int32_t)jj;
                    kept_out[written] = (uint8_t)kept;
                }
                ++written;
            }
        }
    }
    if (fire_count) *fire_count = fires;
    return written;
}

/* ====================================================== projection windows */
/* ProjectionOrder::refresh direction (standard_umap/modern_negatives.rs:208-218,
 * 283-297). */
void orc_projection_direction(orc_rng *rng, int d, float *dir) {
    double norm = 0.0;
    for (int c = 0; c < d; c += 2) {
        double g1, g2;
        orc_gaussian_pair(rng, &g1, &g2);
        dir[c] = (float)g1;
        norm += (double)dir[c] * (double)dir[c];
        if (c + 1 < d) {
            dir[c + 1] = (float)g2;
            norm += (double)dir[c + 1] * (double)dir[c + 1];
        }
    }
    double len = sqrt(norm);
    if (!isfinite(len) || len == 0.0) {
        for (int c = 0; c < d; ++c) dir[c] = 0.0f;
        dir[0] = 1.0f;
    } else {
        for (int c = 0; c < d; ++c) dir[c] = (float)((double)dir[c] / len);
    }
}

/* project_kernel (kernels/standard_umap/modern_layout.rs:259-269). */
void orc_project(const float *x, int64_t rows, int64_t d, int64_t s0, int64_t s1,
                 const float *dir, float *out) {
    for (int64_t i = 0; i < rows; ++i) {
        float sum = 0.0f;
        for (int64_t c = 0; c < d; ++c) {
            float p = x[i * s0 + c * s1] * dir[c];
            sum += p;
        }
        out[i] = sum;
    }
}

/* radix_key_f32 (cuscan-common/src/radix.rs:993-1003). */
uint32_t orc_radix_key_f32(float value) {
    uint32_t bits;
    memcpy(&bits, &value, 4);
    if (bits == 0x80000000u) bits = 0;
    return (bits & 0x80000000u) ? ~bits : (bits ^ 0x80000000u);
}

/* Stable full-key argsort (cuscan-common/src/radix.rs:1-22): LSD radix over
 * the monotone key, 8-bit digits. */
int64_t orc_stable_argsort_f32(const float *values, int64_t n, int32_t *order, float *sorted) {
    for (int64_t i = 0; i < n; ++i)
        if (!isfinite(values[i])) return i;
    uint32_t *keys = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1) * 2);
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    uint32_t *k2 = keys + n;
    for (int64_t i = 0; i < n; ++i) {
        keys[i] = orc_radix_key_f32(values[i]);
        order[i] = (int32_t)i;
    }
    int32_t *src = order, *dst = tmp;
    uint32_t *ks = keys, *kd = k2;
    for (int shift = 0; shift < 32; shift += 8) {
        int64_t cnt[257] = {0};
        for (int64_t i = 0; i < n; ++i) cnt[((ks[i] >> shift) & 255u) + 1]++;
        for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
        for (int64_t i = 0; i < n; ++i) {
            int64_t p = cnt[(ks[i] >> shift) & 255u]++;
            dst[p] = src[i];
            kd[p] = ks[i];
        }
        int32_t *t = src; src = dst; dst = t;
        uint32_t *tk = ks; ks = kd; kd = tk;
    }
    /* four passes: result is back in `order` */
    for (int64_t i = 0; i < n; ++i) sorted[i] = values[order[i]];
    free(keys);
    free(tmp);
    return -1;
}

/* inverse_rank_kernel (:271-277). */
void orc_inverse_ranks(const int32_t *order, int64_t n, int32_t *ranks) {
    for (int64_t j = 0; j < n; ++j) ranks[order[j]] = (int32_t)j;
}

/* lower_bound_kernel (:279-297). */
void orc_lower_bound_ranks(const float *sorted, int64_t n, const float *query, int64_t h,
                           int32_t *ranks) {
    for (int64_t i = 0; i < h; ++i) {
        float value = query[i];
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            int64_t mid = lo + (hi - lo) / 2;
            if (sorted[mid] < value) lo = mid + 1;
            else hi = mid;
        }
        ranks[i] = (int32_t)(lo < n - 1 ? lo : n - 1);
    }
}

/* ======================================================== scale adaptation */
/* resultant_factor / resultant_candidate / resultant_kernel
 * (kernels/standard_umap/modern_layout.rs:301-474), float64. */
static double resultant_factor(const float *emb, int64_t d, const uint32_t *offsets,
                               const uint32_t *indices, int64_t src, int64_t j, double a,
                               double b, double gamma, int exclude) {
    if (exclude) {
        int64_t lo = offsets[src], hi = offsets[src + 1], stop = hi;
        while (lo < hi) {
            int64_t mid = lo + (hi - lo) / 2;
            if ((int64_t)indices[mid] < j) lo = mid + 1;
            else hi = mid;
        }
        int hit = lo < stop && (int64_t)indices[lo] == j;
        if (hit || src == j) return 0.0;
    }
    double q = 0.0;
    for (int64_t c = 0; c < d; ++c) {
        double z = (double)emb[src * d + c] - (double)emb[j * d + c];
        q += z * z;
    }
    double coefficient = 2.0 * gamma * b / ((0.001 + q) * (a * pow(q, b) + 1.0));
    double norm = coefficient * sqrt(q);
    if (norm > 0.0) return coefficient * gamma * tanh(norm / gamma) / norm;
    return 0.0;
}

void orc_resultants(const float *emb, int64_t n, int64_t d, const int32_t *order,
                    const uint32_t *offsets, const uint32_t *indices,
                    const int32_t *sample_ranks, const int32_t *local_offsets,
                    const int32_t *global_ranks, int64_t samples, uint32_t raw_range,
                    float a, float b, float gamma, int exclude, double *out) {
    for (int64_t t = 0; t < 2 * samples; ++t) {
        int local = t >= samples;
        int64_t s = local ? t - samples : t;
        int64_t src = (uint32_t)order[(uint32_t)sample_ranks[s]];
        int64_t js[5];
        for (int k = 0; k < 5; ++k) {
            int64_t rank = (uint32_t)global_ranks[s * 5 + k];
            if (local) {
                int64_t range = raw_range;
                int64_t source_rank = (uint32_t)sample_ranks[s];
                int64_t half = range / 2;
                int64_t first = source_rank > half ? source_rank - half : 0;
                if (first > n - range) first = n - range;
                rank = first + (uint32_t)local_offsets[s * 5 + k];
            }
            js[k] = (uint32_t)order[rank];
        }
        double f[5];
        for (int k = 0; k < 5; ++k)
            f[k] = resultant_factor(emb, d, offsets, indices, src, js[k], (double)a, (double)b,
                                    (double)gamma, exclude);
        double squared = 0.0;
        for (int64_t c = 0; c < d; ++c) {
            double x = (double)emb[src * d + c];
            double sum = f[0] * (x - (double)emb[js[0] * d + c]) +
                         f[1] * (x - (double)emb[js[1] * d + c]) +
                         f[2] * (x - (double)emb[js[2] * d + c]) +
                         f[3] * (x - (double)emb[js[3] * d + c]) +
                         f[4] * (x - (double)emb[js[4] * d + c]);
            squared += sum * sum;
        }
        out[t] = sqrt(squared);
    }
}

static int cmp_double(const void *x, const void *y) {
    double a = *(const double *)x, b = *(const double *)y;
    return a < b ? -1 : (a > b ? 1 : 0);
}

/* negative_force_ratio (umap-core/src/modern_layout.rs:565-600). */
double orc_force_ratio(const double *global, const double *local, int64_t samples) {
    double *logs = (double *)malloc(sizeof(double) * (size_t)(samples ? samples : 1));
    int64_t count = 0;
    for (int64_t s = 0; s < samples; ++s) {
        double g = global[s], l = local[s];
        if (!isfinite(g) || !isfinite(l) || g < 0.0 || l < 0.0) { free(logs); return -1.0; }
        if (g > 0.0 && l > 0.0) logs[count++] = log(g) - log(l);
    }
    if (count == 0) { free(logs); return 1.0; }
    qsort(logs, (size_t)count, sizeof(double), cmp_double);
    double qv[2];
    double qs[2] = {0.1, 0.9};
    for (int k = 0; k < 2; ++k) {
        double position = (double)(count - 1) * qs[k];
        int64_t lower = (int64_t)floor(position), upper = (int64_t)ceil(position);
        qv[k] = logs[lower] + (logs[upper] - logs[lower]) * (position - (double)lower);
    }
    double sum = 0.0;
    for (int64_t s = 0; s < count; ++s) {
        double v = logs[s];
        v = v < qv[0] ? qv[0] : (v > qv[1] ? qv[1] : v);
        sum += v;
    }
    free(logs);
    double ratio = exp(sum / (double)count);
    if (!isfinite(ratio) || ratio <= 0.0) return -1.0;
    return ratio;
}

/* NegativeScaleAdaptation (umap-core/src/modern_layout.rs:604-653). */
void orc_scale_init(orc_scale_state *s, float scale) {
    s->prior = scale;
    s->current = scale;
    s->anchor = 0.0;
    s->has_anchor = 0;
}

float orc_scale_observe(orc_scale_state *s, double ratio) {
    if (s->prior == 0.0) {
        s->anchor = 0.0;
        s->has_anchor = 1;
        return 0.0f;
    }
    if (s->has_anchor) {
        double target = s->anchor * ratio;
        double lo = 0.5 * s->prior, hi = 2.0 * s->prior;
        target = target < lo ? lo : (target > hi ? hi : target);
        double change = log(target / s->current);
        double clo = log(0.9), chi = log(1.1);
        change = change < clo ? clo : (change > chi ? chi : change);
        s->current *= exp(0.1 * change);
    } else {
        s->anchor = s->current / ratio;
        s->has_anchor = 1;
    }
    return (float)s->current;
}

/* ============================================================ epoch driver */
typedef struct {
    orc_rng rng;
    int32_t *proj_order, *ranks;
    float *reference_projection, *query_projection, *sorted;
    float *direction;
} projection_state;

static int projection_refresh(projection_state *P, int64_t h, int64_t n, int64_t d,
                              const float *head, const float *tail, int shared) {
    orc_projection_direction(&P->rng, (int)d, P->direction);
    const float *reference = shared ? head : tail;
    orc_project(reference, n, d, d, 1, P->direction, P->reference_projection);
    if (orc_stable_argsort_f32(P->reference_projection, n, P->proj_order, P->sorted) >= 0)
        return -3;
    if (shared) {
        orc_inverse_ranks(P->proj_order, n, P->ranks);
    } else {
        orc_project(head, h, d, d, 1, P->direction, P->query_projection);
        for (int64_t i = 0; i < h; ++i)
            if (!isfinite(P->query_projection[i])) return -3;
        orc_lower_bound_ranks(P->sorted, n, P->query_projection, h, P->ranks);
    }
    return 0;
}

/* optimize_compact_inner (standard_umap/modern_runtime.rs:1037-1378). */
int orc_optimize_compact(int64_t h, int64_t n, int64_t d, float *head, const float *tail,
                         const uint32_t *offsets, const uint32_t *indices,
                         const double *periods, double *next_positive,
                         const double *negative_periods, double *next_negative,
                         const orc_layout_params *P, int shared, int threads, int stop_epoch,
                         orc_layout_stats *stats) {
    int epochs = P->epochs;
    if (stats) memset(stats, 0, sizeof(*stats));
    if (epochs == 0) return 0;
    if (shared) { n = h; tail = head; }
    int momentum = P->optimizer == 1;
    int64_t Q = P->negative_selection_range;
    int hard = Q < n;
    int alias = shared && momentum && !hard;
    int rc = 0;

    float *updates = (float *)calloc((size_t)(h * d), sizeof(float));
    float *m = momentum ? updates : (float *)calloc((size_t)(h * d), sizeof(float));
    float *v = momentum ? updates : (float *)calloc((size_t)(h * d), sizeof(float));
    int32_t *source = (int32_t *)malloc(sizeof(int32_t) * (size_t)h);
    int32_t *negative = alias ? source : (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int64_t i = 0; i < h; ++i) source[i] = (int32_t)i;
    if (!alias)
        for (int64_t i = 0; i < n; ++i) negative[i] = (int32_t)i;
    int negative_live = !alias;
    int64_t obj_count = 0;
    orc_rng order_rng;
    orc_rng_seed(&order_rng, P->seed);

    projection_state proj;
    memset(&proj, 0, sizeof(proj));
    int projection_ready = 0;
    if (hard) {
        orc_rng_seed(&proj.rng, P->seed);
        proj.proj_order = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
        proj.ranks = (int32_t *)malloc(sizeof(int32_t) * (size_t)(shared ? n : h));
        proj.reference_projection = (float *)malloc(sizeof(float) * (size_t)n);
        proj.query_projection = (float *)malloc(sizeof(float) * (size_t)h);
        proj.sorted = (float *)malloc(sizeof(float) * (size_t)n);
        proj.direction = (float *)malloc(sizeof(float) * (size_t)d);
        projection_ready = h > 1024;
        if (projection_ready) {
            rc = projection_refresh(&proj, h, n, d, head, tail, shared);
            if (stats) stats->refreshes++;
            negative_live = 0;
        }
    }
    int sampling = hard && shared && !momentum && P->adaptation_samples > 0;
    orc_rng force_rng;
    int64_t count = 0;
    int32_t *s_src = NULL, *s_loc = NULL, *s_glob = NULL;
    double *res = NULL;
    if (sampling) {
        orc_rng_seed(&force_rng, 42);
        count = P->adaptation_samples < n ? P->adaptation_samples : n;
        s_src = (int32_t *)malloc(sizeof(int32_t) * (size_t)count);
        s_loc = (int32_t *)malloc(sizeof(int32_t) * (size_t)count * 5);
        s_glob = (int32_t *)malloc(sizeof(int32_t) * (size_t)count * 5);
        res = (double *)malloc(sizeof(double) * (size_t)count * 2);
    }
    orc_scale_state scale;
    orc_scale_init(&scale, orc_resolved_negative_scale((uint64_t)Q, (uint64_t)n, P->has_scale,
                                                       P->scale, P->adaptation_samples > 0));
    uint32_t range = (uint32_t)orc_effective_negative_range((uint64_t)Q, (uint64_t)n);
    double *alpha_s = (double *)malloc(sizeof(double) * (size_t)epochs);
    double *mom_s = (double *)malloc(sizeof(double) * (size_t)epochs);
    if (momentum) orc_momentum_schedule(epochs, P->initial_alpha, P->good, alpha_s, mom_s);
    else orc_adam_schedule(epochs, P->initial_alpha, P->good, alpha_s, mom_s);

    int last = stop_epoch > 0 && stop_epoch < epochs ? stop_epoch : epochs;
    for (int epoch = 0; epoch < last && rc == 0; ++epoch) {
        const int32_t *neg_order = projection_ready ? proj.proj_order : negative;
        const int32_t *ranks = projection_ready ? proj.ranks : source;
        if (sampling && orc_adaptation_due(shared, !momentum, (uint64_t)P->adaptation_samples,
                                           (uint64_t)Q, (uint64_t)n, (uint64_t)epoch,
                                           (uint64_t)epochs)) {
            uint32_t nn = (uint32_t)n, qq = (uint32_t)Q;
            for (int64_t s = 0; s < count; ++s) s_src[s] = (int32_t)orc_rng_range_u32(&force_rng, nn);
            for (int64_t s = 0; s < count * 5; ++s)
                s_loc[s] = (int32_t)orc_rng_range_u32(&force_rng, qq);
            for (int64_t s = 0; s < count * 5; ++s)
                s_glob[s] = (int32_t)orc_rng_range_u32(&force_rng, nn);
            orc_resultants(head, n, d, neg_order, offsets, indices, s_src, s_loc, s_glob, count,
                           qq, P->a, P->b, P->repulsion, P->exclude, res);
            double ratio = orc_force_ratio(res, res + count, count);
            if (ratio <= 0.0) { rc = -3; break; }
            orc_scale_observe(&scale, ratio);
            if (stats) stats->adaptations++;
        }
        orc_epoch_args A;
        memset(&A, 0, sizeof(A));
        A.h = h; A.n = n; A.d = d;
        A.head = head; A.head_s0 = d; A.head_s1 = 1;
        A.tail = tail; A.tail_s0 = d; A.tail_s1 = 1;
        A.offsets = offsets; A.indices = indices;
        A.source_order = source; A.negative_order = neg_order; A.source_ranks = ranks;
        A.periods = periods; A.next_positive = next_positive;
        A.negative_periods = negative_periods; A.next_negative = next_negative;
        A.updates = updates; A.upd_s0 = d; A.upd_s1 = 1;
        A.m = m; A.v = v;
        A.epoch = (uint32_t)epoch;
        A.a = P->a; A.b = P->b; A.gamma = P->repulsion;
        A.scale = (float)scale.current;
        A.negative_range = range;
        A.alpha = (float)alpha_s[epoch];
        A.mu = (float)mom_s[epoch];
        A.correction1 = (float)(1.0 - pow((double)0.9f, (double)epoch + 1.0));
        A.correction2 = (float)(1.0 - pow((double)0.99f, (double)epoch + 1.0));
        A.momentum = momentum;
        A.shared = shared;
        A.negatives = P->negative_sample_rate != 0;
        A.exclude = P->exclude;
        if (stats) {
            int64_t fires = 0;
            int64_t draws = orc_trace_negatives(&A, 0, NULL, NULL, NULL, &fires);
            stats->firing_edges += fires;
            stats->negative_samples += draws;
        }
        orc_modern_epoch(&A, threads);
        if (P->obj_every > 0 && P->obj_out &&
            ((epoch + 1) % P->obj_every == 0 || epoch + 1 == epochs))
            P->obj_out[obj_count++] = orc_cross_entropy(head, h, shared ? NULL : tail, d, offsets,
                                                        indices, P->obj_weights, P->a, P->b);
        if (hard && epoch == 0) negative_live = 0;
        /* prepare_next: the source order first, then the live negative order,
         * from the same generator (modern_orders.rs:544-631). */
        orc_rng_shuffle_i32(&order_rng, source, (size_t)h);
        if (negative_live) orc_rng_shuffle_i32(&order_rng, negative, (size_t)n);
        if (hard) {
            rc = projection_refresh(&proj, h, n, d, head, tail, shared);
            if (stats) stats->refreshes++;
            projection_ready = 1;
        }
        if (stats) stats->epochs_run++;
    }
    if (stats) stats->final_scale = (float)scale.current;
    free(alpha_s); free(mom_s);
    free(s_src); free(s_loc); free(s_glob); free(res);
    if (hard) {
        free(proj.proj_order); free(proj.ranks); free(proj.reference_projection);
        free(proj.query_projection); free(proj.sorted); free(proj.direction);
    }
    if (!alias) free(negative);
    free(source);
    if (!momentum) { free(m); free(v); }
    free(updates);
    return rc;
}

/* ==================================================== graph construction */
/* exact_kernel (cuscan-common/src/kernels/neighbors.rs:17-76): f32
 * sequential squared distance, insertion top-k with lower-index ties,
 * sqrt at the end, sentinel id nr / f32::MAX. */
void orc_exact_knn(const float *query, int64_t nq, const float *ref, int64_t nr, int64_t d,
                   int64_t k, int exclude_self, int64_t query_offset, int threads,
                   int32_t *idx, float *dist) {
    (void)threads;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
    for (int64_t row = 0; row < nq; ++row) {
        int32_t *I = idx + row * k;
        float *D = dist + row * k;
        for (int64_t s = 0; s < k; ++s) {
            I[s] = (int32_t)nr;
            D[s] = FLT_MAX;
        }
        const float *qv = query + row * d;
        for (int64_t r = 0; r < nr; ++r) {
            if (exclude_self && r == query_offset + row) continue;
            const float *rv = ref + r * d;
            float distance = 0.0f;
            for (int64_t f = 0; f < d; ++f) {
                float delta = qv[f] - rv[f];
                float dd = delta * delta;
                distance += dd;
            }
            int64_t worst = k - 1;
            if (distance < D[worst] || (distance == D[worst] && r < I[worst])) {
                int64_t slot = k - 1;
                while (slot > 0) {
                    int64_t prev = slot - 1;
                    if (distance < D[prev] || (distance == D[prev] && r < I[prev])) {
                        I[slot] = I[prev];
                        D[slot] = D[prev];
                        slot -= 1;
                    } else {
                        break;
                    }
                }
                I[slot] = (int32_t)r;
                D[slot] = distance;
            }
        }
        for (int64_t s = 0; s < k; ++s) D[s] = sqrtf(D[s]);
    }
}

/* smooth_row (kernels/standard_umap/graph.rs:117-242). */
static void smooth_row(const float *dist, int64_t k, float global_mean, uint32_t floor_conn,
                       float fraction, uint32_t minimum_positive, int interpolate,
                       float target, float *rho_out, float *sigma_out) {
    int64_t positive_count = 0;
    float positive_max = 0.0f;
    for (int64_t s = 0; s < k; ++s) {
        float x = dist[s];
        if (x > 0.0f) {
            positive_count++;
            if (x > positive_max) positive_max = x;
        }
    }
    float rho = 0.0f;
    if (positive_count > 0) {
        if (floor_conn == 0) {
            float first = positive_max;
            for (int64_t s = 0; s < k; ++s)
                if (dist[s] > 0.0f && dist[s] < first) first = dist[s];
            rho = first * fraction;
        } else {
            int64_t lo_rank = (int64_t)floor_conn - 1, hi_rank = floor_conn;
            float before = positive_max, after = positive_max;
            for (int64_t cs = 0; cs < k; ++cs) {
                float cand = dist[cs];
                if (!(cand > 0.0f)) continue;
                int64_t less = 0, equal = 0;
                for (int64_t o = 0; o < k; ++o) {
                    float other = dist[o];
                    if (other > 0.0f && other < cand) less++;
                    else if (other > 0.0f && other == cand) equal++;
                }
                if (less <= lo_rank && lo_rank < less + equal) before = cand;
                if (less <= hi_rank && hi_rank < less + equal) after = cand;
            }
            if (interpolate) {
                float diff = after - before;
                float t = diff * fraction;
                rho = before + t;
            } else {
                rho = before;
            }
        }
        if (positive_count < (int64_t)minimum_positive) rho = positive_max;
    }
    float row_sum = 0.0f;
    for (int64_t s = 0; s < k; ++s) row_sum += dist[s];
    float floor_mean = rho > 0.0f ? row_sum / (float)(uint32_t)k : global_mean;
    float lo = 0.0f, hi = 3.4028235e38f, mid = 1.0f;
    for (int it = 0; it < 64; ++it) {
        float sum = 0.0f;
        for (int64_t s = 1; s < k; ++s) {
            float delta = dist[s] - rho;
            float safe_mid = mid == 0.0f ? 1.0f : mid;
            float clamped = delta < 0.0f ? 0.0f : delta;
            float value = expf(-clamped / safe_mid);
            if (delta <= 0.0f) value = 1.0f;
            sum += value;
        }
        float error = fabsf(sum - target);
        if (error < 1e-5f) break;
        if (sum > target) {
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
    float floor = floor_mean * 1e-3f;
    if (mid < floor) mid = floor;
    *rho_out = rho;
    *sigma_out = mid;
}

/* fuzzy_graph (kernels/standard_umap/graph.rs:21-43, 252-564). The global
 * mean (which only feeds the sigma floor) is summed in float64 and rounded;
 * the reference leaves its f32 summation order to the backend. */
long orc_fuzzy_graph(const int32_t *indices, const float *distances, int64_t n, int64_t k,
                     float local_connectivity, float bandwidth, float mix, int bipartite,
                     int64_t n_reference, int32_t *heads, int32_t *tails, float *weights,
                     float *rho_out, float *sigma_out, float *directed) {
    int64_t capacity = 2 * n * k;
    uint64_t floor_c = (uint64_t)local_connectivity;
    float fraction = local_connectivity - (float)floor_c;
    uint64_t minimum_positive = floor_c + (local_connectivity > (float)floor_c ? 1 : 0);
    uint64_t above = (uint64_t)k + 1;
    uint32_t fc = (uint32_t)(floor_c < above ? floor_c : above);
    uint32_t mp = (uint32_t)(minimum_positive < above ? minimum_positive : above);
    int interpolate = fraction > 1e-5f;
    float target = log2f((float)k) * bandwidth;
    double total = 0.0;
    for (int64_t i = 0; i < n * k; ++i) total += distances[i];
    float global_mean = (float)(total / (double)(n * k));
    int invalid = 0;
    for (int64_t row = 0; row < n; ++row) {
        float rho, sigma;
        smooth_row(distances + row * k, k, global_mean, fc, fraction, mp, interpolate, target,
                   &rho, &sigma);
        rho_out[row] = rho;
        sigma_out[row] = sigma;
        for (int64_t s = 0; s < k; ++s) {
            int32_t index = indices[row * k + s];
            float dd = distances[row * k + s];
            float delta = dd - rho;
            float safe_sigma = sigma == 0.0f ? 1.0f : sigma;
            float clamped = delta < 0.0f ? 0.0f : delta;
            float value = expf(-clamped / safe_sigma);
            if (dd <= rho || sigma == 0.0f) value = 1.0f;
            if (index < 0 || index >= n_reference || (!bipartite && index == row)) value = 0.0f;
            directed[row * k + s] = value;
            if (index < -1 || index >= n_reference) invalid = 1;
        }
    }
    int64_t nodes = bipartite ? n + n_reference : n;
    int64_t *counts = (int64_t *)calloc((size_t)nodes + 1, sizeof(int64_t));
    /* evaluate_edge, two passes: count then emit in per-head tail order. */
    typedef struct { int32_t h1, t1, h2, t2; float w; int dirs; } emitted;
    emitted *ev = (emitted *)malloc(sizeof(emitted) * (size_t)(n * k > 0 ? n * k : 1));
    for (int64_t row = 0; row < n; ++row) {
        for (int64_t slot = 0; slot < k; ++slot) {
            emitted *E = &ev[row * k + slot];
            E->dirs = 0;
            int32_t tail = indices[row * k + slot];
            if (!(tail >= 0 && tail < n_reference && (bipartite || tail != row))) continue;
            int first = 1;
            float forward = 0.0f;
            for (int64_t o = 0; o < k; ++o)
                if (indices[row * k + o] == tail) {
                    if (o < slot) first = 0;
                    forward += directed[row * k + o];
                }
            if (!first) continue;
            float reverse = 0.0f;
            int has_reverse = 0;
            if (!bipartite) {
                for (int64_t rs = 0; rs < k; ++rs)
                    if (indices[(int64_t)tail * k + rs] == (int32_t)row) {
                        reverse += directed[(int64_t)tail * k + rs];
                        has_reverse = 1;
                    }
            }
            if (!(!has_reverse || row <= tail)) continue;
            float fuzzy;
            if (bipartite) {
                fuzzy = forward;
            } else {
                float inter = forward * reverse;
                float uni = forward + reverse;
                uni = uni - inter;
                float p1 = mix * uni;
                float p2 = (1.0f - mix) * inter;
                fuzzy = p1 + p2;
            }
            if (!(fuzzy > 0.0f)) continue;
            int32_t r32 = (int32_t)row;
            int32_t low = (bipartite || r32 < tail) ? r32 : tail;
            int32_t high = (bipartite || r32 < tail) ? tail : r32;
            E->h1 = low; E->t1 = high; E->h2 = high; E->t2 = low;
            E->w = fuzzy;
            E->dirs = bipartite ? 1 : 2;
            counts[low + 1]++;
            if (E->dirs > 1) counts[high + 1]++;
        }
    }
    for (int64_t i = 0; i < nodes; ++i) counts[i + 1] += counts[i];
    int64_t used = counts[nodes];
    int64_t *cursor = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nodes ? nodes : 1));
    for (int64_t i = 0; i < nodes; ++i) cursor[i] = counts[i];
    for (int64_t i = 0; i < capacity; ++i) { heads[i] = 0; tails[i] = 0; weights[i] = 0.0f; }
    for (int64_t e = 0; e < n * k; ++e) {
        emitted *E = &ev[e];
        if (!E->dirs) continue;
        int64_t p = cursor[E->h1]++;
        heads[p] = E->h1; tails[p] = E->t1; weights[p] = E->w;
        if (E->dirs > 1) {
            p = cursor[E->h2]++;
            heads[p] = E->h2; tails[p] = E->t2; weights[p] = E->w;
        }
    }
    /* per-head insertion sort by tail (tails unique per head) */
    for (int64_t hd = 0; hd < nodes; ++hd) {
        for (int64_t pos = counts[hd] + 1; pos < counts[hd + 1]; ++pos) {
            int32_t hv = heads[pos], tv = tails[pos];
            float wv = weights[pos];
            int64_t ins = pos;
            while (ins > counts[hd] && tails[ins - 1] >= tv) {
                heads[ins] = heads[ins - 1]; tails[ins] = tails[ins - 1];
                weights[ins] = weights[ins - 1];
                ins--;
            }
            heads[ins] = hv; tails[ins] = tv; weights[ins] = wv;
        }
    }
    if (invalid && capacity > 0) { heads[0] = -2; tails[0] = -2; weights[0] = -1.0f; }
    free(ev);
    free(cursor);
    free(counts);
    return invalid ? -1 : (long)used;
}

/* fit_curve (umap-core/src/config.rs:164-330): Levenberg-Marquardt on
 * (log a, log b) over 300 samples of [0, 3*spread]. */
static double curve_value(double x, double a, double b) {
    return x == 0.0 ? 1.0 : 1.0 / (1.0 + a * exp(2.0 * b * log(x)));
}

int orc_fit_curve(float min_dist, float spread, float *a_out, float *b_out) {
    double xs[300], ts[300];
    for (int i = 0; i < 300; ++i) {
        double x = 3.0 * (double)spread * (double)i / 299.0;
        xs[i] = x;
        ts[i] = x < (double)min_dist ? 1.0 : exp(-(x - (double)min_dist) / (double)spread);
    }
    double la = 0.0, lb = 0.0, damping = 1.0e-3;
#define OBJ(LA, LB, OUT)                                          \
    do {                                                          \
        double aa = exp(LA), bb = exp(LB), sum = 0.0;             \
        for (int i = 0; i < 300; ++i) {                           \
            double r = curve_value(xs[i], aa, bb) - ts[i];        \
            sum += r * r;                                         \
        }                                                         \
        OUT = sum;                                                \
    } while (0)
    double objective;
    OBJ(la, lb, objective);
    int converged = 0;
    for (int iter = 0; iter < 500; ++iter) {
        double a = exp(la), b = exp(lb);
        double haa = 0, hab = 0, hbb = 0, ga = 0, gb = 0;
        for (int i = 0; i < 300; ++i) {
            double x = xs[i];
            double r = curve_value(x, a, b) - ts[i];
            double ja = 0.0, jb = 0.0;
            if (x != 0.0) {
                double powered = exp(2.0 * b * log(x));
                double scaled = a * powered;
                double dv = -scaled / ((1.0 + scaled) * (1.0 + scaled));
                ja = dv;
                jb = dv * 2.0 * b * log(x);
            }
            haa += ja * ja; hab += ja * jb; hbb += jb * jb; ga += ja * r; gb += jb * r;
        }
        double gn = fabs(ga) > fabs(gb) ? fabs(ga) : fabs(gb);
        if (gn <= 1.0e-8) { converged = 1; break; }
        double daa = haa + damping * (haa > 1.0 ? haa : 1.0);
        double dbb = hbb + damping * (hbb > 1.0 ? hbb : 1.0);
        double det = daa * dbb - hab * hab;
        if (!isfinite(det) || det <= 0.0) {
            damping *= 10.0;
            if (!isfinite(damping) || damping > 1.0e30) break;
            continue;
        }
        double da = (dbb * ga - hab * gb) / det;
        double db = (daa * gb - hab * ga) / det;
        if (!isfinite(da) || !isfinite(db)) return -3;
        double ca = la - da, cb = lb - db, cobj;
        OBJ(ca, cb, cobj);
        if (cobj < objective) {
            la = ca; lb = cb; objective = cobj;
            damping = damping / 3.0 > 1.0e-15 ? damping / 3.0 : 1.0e-15;
            if ((fabs(da) > fabs(db) ? fabs(da) : fabs(db)) <= 1.0e-11) { converged = 1; break; }
        } else {
            damping *= 10.0;
            if (!isfinite(damping) || damping > 1.0e30) break;
        }
    }
#undef OBJ
    double a = exp(la), b = exp(lb);
    if (!converged || !isfinite(a) || !isfinite(b) || a <= 0.0 || b <= 0.0) return -3;
    *a_out = (float)a;
    *b_out = (float)b;
    return 0;
}

/* ================================================= initialisation (fit) */
static int cmp_dbl(const void *x, const void *y) {
    double a = *(const double *)x, b = *(const double *)y;
    return a < b ? -1 : (a > b ? 1 : 0);
}

static double quantile_sorted(const double *s, int64_t n, double p) {
    double h = (double)(n - 1) * p;
    int64_t lo = (int64_t)floor(h), hi = (int64_t)ceil(h);
    return s[lo] + (s[hi] - s[lo]) * (h - (double)lo);
}

/* random_embedding (umap-core/src/init.rs:1519-1530) + rescale
 * (init_rescale.rs:77-203), isotropic (mode 0) or per axis (mode 1). */
int orc_random_init(int64_t n, int64_t d, uint64_t seed, float repulsion, float factor,
                    int mode, float *out) {
    if (n == 1) {
        for (int64_t c = 0; c < d; ++c) out[c] = 0.0f;
        return 0;
    }
    orc_rng rng;
    orc_rng_seed(&rng, seed);
    for (int64_t i = 0; i < n * d; ++i) out[i] = orc_rng_range_f32(&rng, -10.0f, 10.0f);
    for (int64_t c = 0; c < d; ++c) {
        float sum = 0.0f;
        for (int64_t r = 0; r < n; ++r) sum += out[r * d + c];
        float mean = sum / (float)n;
        for (int64_t r = 0; r < n; ++r) out[r * d + c] -= mean;
    }
    double *scratch = (double *)malloc(sizeof(double) * (size_t)n);
    double *spans = (double *)malloc(sizeof(double) * (size_t)d);
    for (int64_t c = 0; c < d; ++c) {
        for (int64_t r = 0; r < n; ++r) scratch[r] = out[r * d + c];
        qsort(scratch, (size_t)n, sizeof(double), cmp_dbl);
        double span = quantile_sorted(scratch, n, 0.95) - quantile_sorted(scratch, n, 0.05);
        spans[c] = span == 0.0 ? 1.0 : span;
    }
    double target = log10((double)n) * 3.0 * log1p((double)repulsion) / 0.6931471805599453 *
                    (double)factor;
    if (mode == 0) {
        double logs = 0.0;
        for (int64_t c = 0; c < d; ++c) logs += log(spans[c]);
        float mult = (float)(target / exp(logs / (double)d));
        for (int64_t i = 0; i < n * d; ++i) out[i] *= mult;
    } else {
        for (int64_t c = 0; c < d; ++c) {
            float mult = (float)(target / spans[c]);
            for (int64_t r = 0; r < n; ++r) out[r * d + c] *= mult;
        }
    }
    free(scratch);
    free(spans);
    return 0;
}

/* weighted_initialization (crates/umap/src/standard_umap/estimator.rs:
 * 2385-2481). Returns -1 (and stops) on a row without positive membership. */
int orc_transform_init(const int32_t *indices, const float *directed, int64_t q, int64_t k,
                       const float *embedding, int64_t d, float *out) {
    for (int64_t row = 0; row < q; ++row) {
        const int32_t *ids = indices + row * k;
        const float *w = directed + row * k;
        int64_t selected = -1;
        int positive = 0;
        for (int64_t s = 0; s < k; ++s) {
            int32_t id = ids[s];
            if (id < 0) continue;
            int seen = 0;
            for (int64_t o = 0; o < s; ++o) seen |= ids[o] == id;
            if (seen) continue;
            float weight = 0.0f;
            for (int64_t o = 0; o < k; ++o)
                if (ids[o] == id) weight += w[o];
            positive |= weight > 0.0f;
            if (weight == 1.0f && (selected < 0 || id < selected)) selected = id;
        }
        if (!positive) return -1;
        if (selected >= 0) {
            for (int64_t c = 0; c < d; ++c) out[row * d + c] = embedding[selected * d + c];
            continue;
        }
        float sum = 0.0f;
        for (int64_t s = 0; s < k; ++s) sum += w[s];
        float div = sum == 0.0f ? 1.0f : sum;
        for (int64_t c = 0; c < d; ++c) out[row * d + c] = 0.0f;
        for (int64_t s = 0; s < k; ++s) {
            float nw = w[s] / div;
            for (int64_t c = 0; c < d; ++c) {
                float prod = embedding[(int64_t)ids[s] * d + c] * nw;
                out[row * d + c] = out[row * d + c] + prod;
            }
        }
    }
    return 0;
}


/* ------------------------------------------------------------------------
 * Compatibility (classic) SGD.
 * Follows crates/umap/src/kernels/standard_umap/layout.rs:68-336 (clock
 * pass :68-100, vertex pass :230-336) and the driver
 * crates/umap/src/standard_umap/estimator.rs:2753-2982.
 * ------------------------------------------------------------------------ */

/* The three Tausworthe components of layout.rs:291-303 on signed 64-bit
 * state: ((s & keep) << l) & 0xffffffff, xor ((((s << m) & 0xffffffff) ^ s)
 * >> r) with an arithmetic right shift. */
static int64_t orc_taus_component(int64_t s, uint64_t keep, int l, int m, int r) {
    uint64_t u = (uint64_t)s;
    uint64_t left = ((u & keep) << l) & 0xffffffffull;
    uint64_t mixed = ((u << m) & 0xffffffffull) ^ u;
    int64_t right = (int64_t)mixed;
    /* arithmetic shift written without relying on implementation-defined >> */
    right = right < 0 ? ~(~right >> r) : right >> r;
    return (int64_t)left ^ right;
}

int64_t orc_taus_next(int64_t st[3]) {
    st[0] = orc_taus_component(st[0], 4294967294ull, 12, 13, 19);
    st[1] = orc_taus_component(st[1], 4294967288ull, 4, 2, 25);
    st[2] = orc_taus_component(st[2], 4294967280ull, 17, 3, 11);
    return (int64_t)(int32_t)(uint32_t)(uint64_t)(st[0] ^ st[1] ^ st[2]);
}

static float orc_classic_dist(const float *head, const float *other, int64_t d) {
    float q = 0.0f;
    for (int64_t c = 0; c < d; ++c) {
        float diff = head[c] - other[c];
        q = q + diff * diff;
    }
    return q;
}

static void orc_classic_move(float *head, float *other, int64_t d, float coef, float alpha,
                             int move_other) {
    for (int64_t c = 0; c < d; ++c) {
        float g = coef * (head[c] - other[c]);
        if (g < -4.0f) g = -4.0f;
        if (g > 4.0f) g = 4.0f;
        float upd = g * alpha;
        head[c] = head[c] + upd;
        if (move_other) other[c] = other[c] - upd;
    }
}

void orc_classic_epoch(int64_t h, int64_t n, int64_t d, float *head, const float *tail,
                       const int32_t *tails, const float *weights, int64_t E,
                       const double *per, double *npos, const double *nper, double *nneg,
                       int64_t *rng, const int32_t *head_offsets, const int32_t *head_edges,
                       uint32_t epoch, float a, float b, float gamma, float alpha, float rate,
                       int shared) {
    int32_t *mask = (int32_t *)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
    double now = (double)epoch;
    for (int64_t e = 0; e < E; ++e) {
        int32_t fired = -1;
        if (weights[e] > 0.0f && npos[e] <= now) {
            fired = 0;
            npos[e] = npos[e] + per[e];
            if (rate > 0.0f) {
                int32_t count = (int32_t)((now - nneg[e]) / nper[e]);
                if (count > 0) fired = count;
                nneg[e] = nneg[e] + (double)count * nper[e];
            }
        }
        mask[e] = fired;
    }
    float *tail_rows = shared ? head : (float *)tail;
    for (int64_t v = 0; v < h; ++v) {
        int64_t *st = rng + 3 * v;
        for (int32_t slot = head_offsets[v]; slot < head_offsets[v + 1]; ++slot) {
            int32_t e = head_edges[slot];
            if (mask[e] < 0) continue;
            int64_t t = tails[e];
            if (t < n && (!shared || t != v)) {
                float q = orc_classic_dist(head + v * d, tail_rows + t * d, d);
                if (q > 0.0f) {
                    float coef = (-2.0f * a * b) * powf(q, b - 1.0f) / (a * powf(q, b) + 1.0f);
                    orc_classic_move(head + v * d, tail_rows + t * d, d, coef, alpha, shared);
                }
            }
            for (int32_t k = 0; k < mask[e]; ++k) {
                int32_t draw = (int32_t)orc_taus_next(st);
                int32_t rem = draw % (int32_t)n;
                if (rem < 0) rem += (int32_t)n;
                if (shared && rem == v) continue;
                float q = orc_classic_dist(head + v * d, tail_rows + (int64_t)rem * d, d);
                if (q > 0.0f) {
                    float coef = (2.0f * gamma * b) / ((0.001f + q) * (a * powf(q, b) + 1.0f));
                    orc_classic_move(head + v * d, tail_rows + (int64_t)rem * d, d, coef, alpha, 0);
                }
            }
        }
    }
    free(mask);
}

int orc_classic_optimize(int64_t h, int64_t n, int64_t d, float *head, const float *tail,
                         const int32_t *heads, const int32_t *tails, const float *weights,
                         int64_t E, int32_t n_epochs, int32_t default_epochs, float a, float b,
                         float learning_rate, float gamma, int32_t rate, uint64_t seed,
                         int shared, double *npos_out, double *nneg_out, int64_t *rng_out) {
    if (shared) n = h;
    size_t Ec = (size_t)(E > 0 ? E : 1);
    float *pruned = (float *)malloc(sizeof(float) * Ec);
    double *per = (double *)malloc(sizeof(double) * Ec), *npos = (double *)malloc(sizeof(double) * Ec);
    double *nper = (double *)malloc(sizeof(double) * Ec), *nneg = (double *)malloc(sizeof(double) * Ec);
    int32_t *offsets = (int32_t *)calloc((size_t)h + 1, sizeof(int32_t));
    int32_t *ids = (int32_t *)malloc(sizeof(int32_t) * Ec);
    int64_t *rng = (int64_t *)malloc(sizeof(int64_t) * (size_t)(3 * h));
    /* prepare_schedule (estimator.rs:2753-2798) */
    float wmax = 0.0f;
    for (int64_t e = 0; e < E; ++e) wmax = weights[e] > wmax ? weights[e] : wmax;
    float threshold = wmax / (float)(n_epochs > 10 ? n_epochs : default_epochs);
    double r = (double)(rate > 1 ? rate : 1);
    for (int64_t e = 0; e < E; ++e) {
        pruned[e] = weights[e] < threshold ? 0.0f : weights[e];
        double safe = pruned[e] == 0.0f ? 1.0 : (double)pruned[e];
        per[e] = (double)wmax / safe;
        npos[e] = per[e];
        nper[e] = per[e] / r;
        nneg[e] = nper[e];
        if (pruned[e] > 0.0f) offsets[heads[e] + 1] += 1;
    }
    /* build_head_csr (estimator.rs:3080-3131): edge ids grouped by head */
    for (int64_t i = 0; i < h; ++i) offsets[i + 1] += offsets[i];
    {
        int32_t *cursor = (int32_t *)malloc(sizeof(int32_t) * (size_t)(h > 0 ? h : 1));
        memcpy(cursor, offsets, sizeof(int32_t) * (size_t)h);
        for (int64_t e = 0; e < E; ++e)
            if (pruned[e] > 0.0f) ids[cursor[heads[e]]++] = (int32_t)e;
        free(cursor);
    }
    /* optimizer_rng_from_first_coordinates (estimator.rs:2800-2829):
     * StdRng(seed) random_range(i32::MIN..i32::MAX) three times, plus the
     * bits of each row's first coordinate widened to f64 */
    orc_rng g;
    orc_rng_seed(&g, seed);
    int64_t base[3];
    for (int k = 0; k < 3; ++k)
        base[k] = (int64_t)(int32_t)(0x80000000u + orc_rng_range_u32(&g, 0xFFFFFFFFu));
    for (int64_t i = 0; i < h; ++i) {
        double first = (double)head[i * d];
        uint64_t bits;
        memcpy(&bits, &first, sizeof(bits));
        for (int k = 0; k < 3; ++k) rng[3 * i + k] = (int64_t)((uint64_t)base[k] + bits);
    }
    /* optimize (estimator.rs:2868-2982): linear learning-rate decay */
    float alpha = learning_rate;
    for (int32_t t = 0; t < n_epochs; ++t) {
        orc_classic_epoch(h, n, d, head, tail, tails, pruned, E, per, npos, nper, nneg, rng,
                          offsets, ids, (uint32_t)t, a, b, gamma, alpha, (float)rate, shared);
        alpha = learning_rate * (1.0f - (float)t / (float)(n_epochs > 1 ? n_epochs : 1));
    }
    if (npos_out) memcpy(npos_out, npos, sizeof(double) * (size_t)E);
    if (nneg_out) memcpy(nneg_out, nneg, sizeof(double) * (size_t)E);
    if (rng_out) memcpy(rng_out, rng, sizeof(int64_t) * (size_t)(3 * h));
    free(pruned); free(per); free(npos); free(nper); free(nneg); free(offsets); free(ids); free(rng);
    return 0;
}


/* ------------------------------------------------------------------------
 * FP64 objective: sum_e -w ln(max(phi, 1e-4)) - (1 - w) ln(max(1 - phi,
 * 1e-4)), phi = 1 / (1 + a q^b), in row-major edge order with Neumaier
 * compensated summation.
 * ------------------------------------------------------------------------ */
double orc_cross_entropy(const float *head, int64_t h, const float *tail, int64_t d,
                         const uint32_t *offsets, const uint32_t *indices, const float *weights,
                         double a, double b) {
    const float *y = tail ? tail : head;
    double sum = 0.0, comp = 0.0;
    for (int64_t i = 0; i < h; ++i) {
        for (uint32_t e = offsets[i]; e < offsets[i + 1]; ++e) {
            int64_t j = indices[e];
            double q = 0.0;
            for (int64_t c = 0; c < d; ++c) {
                double diff = (double)head[i * d + c] - (double)y[j * d + c];
                q += diff * diff;
            }
            double p = q > 0.0 ? a * pow(q, b) : 0.0;
            double phi = 1.0 / (1.0 + p), rest = p / (1.0 + p), w = (double)weights[e];
            double term = -w * log(phi > 1e-4 ? phi : 1e-4) - (1.0 - w) * log(rest > 1e-4 ? rest : 1e-4);
            double t = sum + term;
            if (fabs(sum) >= fabs(term)) comp += (sum - t) + term;
            else comp += (term - t) + sum;
            sum = t;
        }
    }
    return sum + comp;
}
