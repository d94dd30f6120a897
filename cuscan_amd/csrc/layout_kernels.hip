// Modern-layout device kernels for gfx950 (CDNA4, wave64).
//
// Semantics follow crates/umap/src/kernels/standard_umap/modern_layout.rs.
// The build disables floating-point contraction so that the f64 clocks
// (next_negative += count * period) and every f32 force term round exactly
// as the reference's uncontracted arithmetic does.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>

#include "common.hpp"
#include "layout_kernels.hpp"
#include "wave.hpp"
#include "host_parallel.hpp"

namespace umaphip {

namespace {

constexpr int kThreads = 256;

inline unsigned grid_for(int64_t n, int threads = kThreads) {
    return static_cast<unsigned>((n + threads - 1) / threads);
}

// ------------------------------------------------------------ force terms
// Attraction coefficient (gradient_kernel :166-175), multiplier folded as
// factor * multiplier * (-2ab q^(b-1)) / (1 + a q^b).
__device__ inline float attract_coeff(float q, float a, float b, float fm) {
    float num = -2.0f * a;
    num = num * b;
    num = num * powf(q, b - 1.0f);
    const float top = fm * num;
    const float apb = a * powf(q, b);
    return top / (1.0f + apb);
}

// Repulsion coefficient (repel :91-126); returns 0 when the reference
// skips the contribution.
__device__ inline float repel_coeff(float q, float a, float b, float gamma, float s,
                                    float multiplier) {
    if (!(q > 0.0f && gamma > 0.0f)) return 0.0f;
    float num = s * 2.0f;
    num = num * gamma;
    num = num * b;
    const float den1 = 0.001f + q;
    const float apb = a * powf(q, b);
    const float den = den1 * (1.0f + apb);
    const float coeff = num / den;
    if (!(coeff > 0.0f)) return 0.0f;
    const float norm = sqrtf((coeff * coeff) * q);
    if (!(norm > 0.0f)) return 0.0f;
    const float g16 = gamma < 16.0f ? gamma : 16.0f;
    float clip = g16 * tanhf(norm / gamma);
    clip = clip / norm;
    return (multiplier * coeff) * clip;
}

// Binary search of j in a sorted CSR row, or the source itself when fit
// (excluded, :63-89).
__device__ inline bool excluded(const int32_t *indices, int64_t begin, int64_t stop, int64_t i,
                                int64_t j, bool shared) {
    int64_t lo = begin, hi = stop;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (static_cast<int64_t>(static_cast<uint32_t>(indices[mid])) < j) lo = mid + 1;
        else hi = mid;
    }
    bool hit = lo < stop && static_cast<int64_t>(static_cast<uint32_t>(indices[lo])) == j;
    return hit || (shared && i == j);
}

// Negative candidate enumeration (gradient_kernel :177-250). Produces the
// rank into negative_order of each candidate, in the reference's order.
// e * (epoch + p + 1) mod M is advanced by adding e mod M, which equals the
// reference's u64 product whenever that product does not wrap (E < 2^31,
// epoch < 2^32).
struct NegWalk {
    uint64_t modulus;  // R (large) or N (small)
    uint64_t step;     // e mod modulus
    uint64_t cur;      // e*(epoch+1+p) mod modulus
    uint64_t first;    // window start (large)
    uint64_t left;     // samples (plain) or attempts left (retry)
    uint64_t need;     // accepted still required (retry)
    int mode;          // 0 retry, 1 small plain, 2 large
    __device__ inline void init(uint64_t e, uint32_t epoch, int32_t count, bool small,
                                bool exclude, uint64_t n, uint64_t range, int64_t rank) {
        mode = small ? (exclude ? 0 : 1) : 2;
        modulus = small ? n : range;
        step = e % modulus;
        cur = (step * ((static_cast<uint64_t>(epoch) + 1u) % modulus)) % modulus;
        first = 0;
        if (!small && range < n) {
            const uint64_t rk = static_cast<uint64_t>(static_cast<uint32_t>(rank));
            const uint64_t half = range / 2;
            if (rk > half) first = rk - half;
            if (first > n - range) first = n - range;
        }
        if (mode == 0) {
            left = static_cast<uint64_t>(count) * n;
            need = static_cast<uint64_t>(count);
        } else {
            left = static_cast<uint64_t>(count);
            need = 0;
        }
    }
    // Next candidate rank; false when done. For retry, report acceptance
    // with accept() after checking exclusion.
    __device__ inline bool next(uint64_t &rank) {
        if (mode == 0) {
            if (need == 0 || left == 0) return false;
            rank = cur;
            --left;
            cur += 1;
            if (cur >= modulus) cur -= modulus;
            return true;
        }
        if (left == 0) return false;
        rank = first + cur;
        --left;
        cur += step;
        if (cur >= modulus) cur -= modulus;
        return true;
    }
    __device__ inline void accept() {
        if (mode == 0) --need;
    }
};

// ------------------------------------------------- generic per-slot kernel
__device__ inline float head_c(const RowwiseArgs &A, int64_t row, int64_t c) {
    return A.head[row * A.hs0 + c * A.hs1];
}
__device__ inline float tgt_c(const RowwiseArgs &A, int64_t row, int64_t c) {
    return A.shared ? A.head[row * A.hs0 + c * A.hs1] : A.tail[row * A.ts0 + c * A.ts1];
}

__device__ inline void rowwise_accumulate(const RowwiseArgs &A, int64_t i, int64_t j,
                                          float coeff) {
    for (int64_t c = 0; c < A.d; ++c) {
        const float diff = head_c(A, i, c) - tgt_c(A, j, c);
        float *u = &A.updates[i * A.us0 + c * A.us1];
        *u = *u + coeff * diff;
    }
}

__device__ inline float rowwise_distance(const RowwiseArgs &A, int64_t i, int64_t j) {
    float q = 0.0f;
    for (int64_t c = 0; c < A.d; ++c) {
        const float z = head_c(A, i, c) - tgt_c(A, j, c);
        q += z * z;
    }
    return q;
}

__global__ __launch_bounds__(64) void gradient_rowwise_kernel(RowwiseArgs A, int64_t start,
                                                              int64_t end) {
    const int64_t slot = start + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (slot >= end) return;
    const int64_t i = static_cast<uint32_t>(A.source_order[slot]);
    const bool small = A.h <= 1024;
    const double now = static_cast<double>(A.epoch);
    const int64_t begin = A.offsets[i], stop = A.offsets[i + 1];
    for (int64_t e = begin; e < stop; ++e) {
        const double npos = A.next_positive[e];
        if (!(npos <= now)) continue;
        const int64_t j = static_cast<uint32_t>(A.indices[e]);
        const float q = rowwise_distance(A, i, j);
        if (q > 0.0f) {
            const float fm = (small ? 0.5f : 1.0f) * A.multiplier;
            rowwise_accumulate(A, i, j, attract_coeff(q, A.a, A.b, fm));
        }
        A.next_positive[e] = npos + A.periods[e];
        if (A.negatives) {
            const double nn = A.next_negative[e];
            const double np = A.negative_periods[e];
            const int32_t count = static_cast<int32_t>((now - nn) / np);
            if (count > 0) {
                NegWalk w;
                w.init(static_cast<uint64_t>(e), A.epoch, count, small, A.exclude != 0,
                       static_cast<uint64_t>(A.n), A.range,
                       (!small && A.range < static_cast<uint64_t>(A.n)) ? A.source_ranks[i] : 0);
                uint64_t rank;
                const float s = small ? 1.0f : (A.scale_ptr ? *A.scale_ptr : A.scale);
                while (w.next(rank)) {
                    const int64_t jj = static_cast<uint32_t>(A.negative_order[rank]);
                    if (A.exclude && excluded(A.indices, begin, stop, i, jj, A.shared != 0))
                        continue;
                    w.accept();
                    const float c = repel_coeff(rowwise_distance(A, i, jj), A.a, A.b, A.gamma, s,
                                                A.multiplier);
                    if (c != 0.0f) rowwise_accumulate(A, i, jj, c);
                }
            }
            A.next_negative[e] = nn + static_cast<double>(count) * np;
        }
    }
}

__global__ void debug_negatives_kernel(RowwiseArgs A, int32_t cap, int32_t *count_out,
                                       int32_t *cand_out, uint8_t *kept_out, int32_t *overflow) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= A.h) return;
    const bool small = A.h <= 1024;
    const double now = static_cast<double>(A.epoch);
    const int64_t begin = A.offsets[i], stop = A.offsets[i + 1];
    for (int64_t e = begin; e < stop; ++e) {
        if (!(A.next_positive[e] <= now)) {
            count_out[e] = -1;
            continue;
        }
        int32_t visited = 0;
        if (A.negatives) {
            const int32_t count =
                static_cast<int32_t>((now - A.next_negative[e]) / A.negative_periods[e]);
            if (count > 0) {
                NegWalk w;
                w.init(static_cast<uint64_t>(e), A.epoch, count, small, A.exclude != 0,
                       static_cast<uint64_t>(A.n), A.range,
                       (!small && A.range < static_cast<uint64_t>(A.n)) ? A.source_ranks[i] : 0);
                uint64_t rank;
                while (w.next(rank)) {
                    const int64_t jj = static_cast<uint32_t>(A.negative_order[rank]);
                    const bool keep =
                        !(A.exclude && excluded(A.indices, begin, stop, i, jj, A.shared != 0));
                    if (keep) w.accept();
                    if (visited < cap) {
                        cand_out[e * cap + visited] = static_cast<int32_t>(jj);
                        kept_out[e * cap + visited] = keep ? 1 : 0;
                    } else {
                        atomicExch(overflow, 1);
                    }
                    ++visited;
                }
            }
        }
        count_out[e] = visited;
    }
}

__global__ void adam_kernel(int64_t h, int64_t d, float *head, int64_t hs0, int64_t hs1,
                            float *upd, int64_t us0, int64_t us1, float *m, int64_t ms0,
                            int64_t ms1, float *v, int64_t vs0, int64_t vs1, float alpha,
                            float c1, float c2) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= h * d) return;
    const int64_t row = idx / d, col = idx % d;
    float *u = &upd[row * us0 + col * us1];
    const float g = *u;
    if (g != 0.0f) {
        float *mm = &m[row * ms0 + col * ms1];
        float *vv = &v[row * vs0 + col * vs1];
        const float first = 0.9f * *mm + (1.0f - 0.9f) * g;
        const float second = 0.99f * *vv + ((1.0f - 0.99f) * g) * g;
        *mm = first;
        *vv = second;
        float *x = &head[row * hs0 + col * hs1];
        *x = *x + (alpha * (first / c1)) / (sqrtf(second / c2) + 0.0001f);
    }
    *u = 0.0f;
}

__global__ void momentum_apply_kernel(int64_t d, float *head, int64_t hs0, int64_t hs1,
                                      const float *vel, int64_t vs0, int64_t vs1,
                                      const int32_t *order, int64_t start, int64_t end) {
    const int64_t slot = start + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (slot >= end) return;
    const int64_t row = static_cast<uint32_t>(order[slot]);
    for (int64_t c = 0; c < d; ++c) {
        float *x = &head[row * hs0 + c * hs1];
        *x = *x + vel[row * vs0 + c * vs1];
    }
}

__global__ void momentum_decay_kernel(int64_t h, int64_t d, float *vel, int64_t vs0,
                                      int64_t vs1, float mu) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= h * d) return;
    float *x = &vel[(idx / d) * vs0 + (idx % d) * vs1];
    *x = *x * mu;
}

// ------------------------------------------------------ tiled edge sweep
// One workgroup owns kTileRows consecutive CSR rows and streams their edges
// in chunks of 256 (one lane per edge: coalesced clock/index/period reads).
//   A. Each lane advances its edge's f64 clocks and counts its force terms:
//      the attraction plus one term per negative sample.
//   B. The chunk's terms are laid out in edge order in LDS (a block-wide
//      exclusive scan) and computed term-parallel, so every lane does the
//      same work whether its edge fired or not; each negative term derives
//      its candidate rank e*(epoch+p+1) mod M independently.
//   C. One lane per row adds its row's terms in edge order, i.e. the
//      reference's per-term sequential accumulation order. A chunk with
//      more terms than LDS holds runs B and C in several windows, so the
//      sums do not depend on the tile shape.
// The fused epilogue applies the lazy Adam step into a second position
// buffer (Jacobi double buffer). Deterministic; no global atomics.
constexpr int kTileRows = 64;

__device__ inline float2 ld2(const float *p, int64_t row) {
    return reinterpret_cast<const float2 *>(p)[row];
}

// Hardware reciprocal (1 ulp).
__device__ inline float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// q^b by the hardware base-2 exp/log (relative error a few ulp).
__device__ inline float fast_pow(float q, float b) {
    return __builtin_amdgcn_exp2f(b * __builtin_amdgcn_logf(q));
}

// tanh for x >= 0: odd Taylor polynomial below 1/8 (truncation < 2e-10
// relative), (e^{2x} - 1) / (e^{2x} + 1) above, 1 beyond 9.
__device__ inline float tanh_pos(float x) {
    if (x < 0.125f) {
        const float x2 = x * x;
        return x * (1.0f + x2 * (-0.33333334f + x2 * (0.13333334f + x2 * -0.053968254f)));
    }
    if (x > 9.0f) return 1.0f;
    const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
    return (e - 1.0f) * rcpf(e + 1.0f);
}

// Attraction factor (coefficient of x_i - y): fm * -2ab q^(b-1) / (1 + a q^b).
__device__ inline float attract_fast(float q, float a, float b, float fm) {
    const float pb = fast_pow(q, b);
    return fm * (-2.0f * a * b) * (pb * rcpf(q)) * rcpf(1.0f + a * pb);
}

// Repulsion factor: with c = 2*gamma*b*s / ((0.001 + q)(1 + a q^b)) and
// norm = c*sqrt(q), c * min(gamma,16) * tanh(norm/gamma) / norm equals
// min(gamma,16) * tanh(c*sqrt(q)/gamma) / sqrt(q).
__device__ inline float repel_fast(float q, float a, float b, float gamma, float s) {
    if (!(q > 0.0f && gamma > 0.0f)) return 0.0f;
    const float c = s * 2.0f * gamma * b * rcpf((0.001f + q) * (1.0f + a * fast_pow(q, b)));
    if (!(c > 0.0f)) return 0.0f;
    const float rq = __builtin_amdgcn_rsqf(q);
    const float g16 = gamma < 16.0f ? gamma : 16.0f;
    return g16 * tanh_pos(c * q * rq * rcpf(gamma)) * rq;
}

// Either force coefficient with one q^b and one reciprocal, for sweeps whose
// waves mix attraction and repulsion terms: attraction
// k_att q^b / (q (1 + a q^b)) with k_att = -2ab; repulsion as repel_fast with
// k_rep = 2 s gamma b, inv_gamma = 1/gamma and g16 = min(gamma, 16). q > 0.
__device__ inline float unit_coef(float q, bool att, float a, float b, float k_att, float k_rep,
                                  float inv_gamma, float g16) {
    const float pb = fast_pow(q, b);
    const float r = rcpf((att ? q : 0.001f + q) * (1.0f + a * pb));
    if (att) return k_att * pb * r;
    const float c = k_rep * r;
    if (!(c > 0.0f)) return 0.0f;
    const float rq = __builtin_amdgcn_rsqf(q);
    return g16 * tanh_pos(c * q * rq * inv_gamma) * rq;
}

// x mod m through a double-precision quotient estimate. Precondition:
// 0 < m < 2^31 and x < m * 2^31 (every caller passes x < m^2; m is the
// negative range, at most N <= INT32_MAX). Then x / m < 2^31, the estimate
// x * RN(1/m) is within 2^-20 of x / m, so its truncation q is the true
// quotient or one off, x - q m lies in [-m, 2m) and one correction step each
// way finishes. The steps are written without loops: a loop here lets the
// compiler replace its trip count by a 64-bit division.
__device__ inline uint64_t mod_u64(uint64_t x, uint64_t m, double inv_m) {
    const uint64_t q = static_cast<uint32_t>(static_cast<double>(x) * inv_m);
    int64_t r = static_cast<int64_t>(x - q * m);
    r += r < 0 ? static_cast<int64_t>(m) : 0;
    r -= r >= static_cast<int64_t>(m) ? static_cast<int64_t>(m) : 0;
    return static_cast<uint64_t>(r);
}

// x mod m for an integer-valued double 0 <= x < 2^53 with x / m < 2^40 and
// m >= 1 an integer: trunc(x * RN(1/m)) is floor(x / m) or one off, the
// remainder x - q m is an integer in [-m, 2m) that the fma yields exactly,
// and one correction each way finishes.
__device__ inline uint32_t mod_f64(double x, double m, double inv_m) {
    const double q = trunc(x * inv_m);
    double r = fma(-q, m, x);
    r = r < 0.0 ? r + m : r;
    r = r >= m ? r - m : r;
    return static_cast<uint32_t>(r);
}

// Largest f for which e_mod * f (e_mod < m) stays below 2^53, less a margin
// of two for the rounding of inv_m: products up to it are exact in f64.
__device__ inline uint64_t exact_product_limit(double inv_m) {
    const double lim = 9007199254740992.0 * inv_m;
    return lim > 4.0 ? static_cast<uint64_t>(lim) - 2u : 0u;
}

// Candidate offset of the k-th (k >= 1) negative an edge draws at epoch t,
// e (t + k) mod M (modern_layout.rs:433-514), from em = e mod M and
// t1 = (t + 1) mod M. With f = t1 + k - 1 the value is em * f mod M; when
// that product is exact in f64 (f <= limit) one f64 reduction gives it.
// Otherwise e (t + 1) mod M goes through 64-bit integers and the remaining
// k - 1 steps of em are added in f64 (exact for k < 2^22).
__device__ inline uint32_t negative_offset(uint32_t em, uint32_t t1, int k, uint64_t modulus,
                                           double m_d, double inv_m, uint64_t limit) {
    const uint64_t f = static_cast<uint64_t>(t1) + static_cast<uint64_t>(k - 1);
    if (f <= limit)
        return mod_f64(static_cast<double>(em) * static_cast<double>(static_cast<uint32_t>(f)),
                       m_d, inv_m);
    const uint64_t s0 = mod_u64(static_cast<uint64_t>(em) * t1, modulus, inv_m);
    const double x = static_cast<double>(static_cast<uint32_t>(s0)) +
                     static_cast<double>(k - 1) * static_cast<double>(em);
    return mod_f64(x, m_d, inv_m);
}

template <int WAVES>
__device__ inline int block_exclusive_scan(int v, int tid, int *s_wave, int &total) {
    const int lane = tid & 63, wave = tid >> 6;
    int incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int up = __shfl_up(incl, off, 64);
        if (lane >= off) incl += up;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; ++w) before += s_wave[w];
    total = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) total += s_wave[w];
    return before + incl - v;
}

// ROWS rows per tile, THREADS lanes per workgroup (chunks of THREADS edges).
template <int ROWS, int THREADS>
__device__ inline void sweep_tile(const SweepArgs &A, int64_t tile) {
    constexpr int TCAP = 8 * THREADS;
    constexpr int WAVES = THREADS / 64;
    static_assert(ROWS <= 64 && ROWS <= THREADS && THREADS % 64 == 0 && THREADS <= 256,
                  "tile shape");
    __shared__ int32_t s_off[ROWS + 1];   // tile-local edge prefix over the tile's rows
    __shared__ int32_t s_seg[ROWS];       // CSR start of each tile row
    __shared__ int32_t s_rowid[ROWS];
    __shared__ float2 s_pos[ROWS];
    __shared__ uint32_t s_first[ROWS];
    __shared__ int8_t s_active[ROWS];
    __shared__ float2 s_terms[TCAP];
    __shared__ uint8_t s_owner[TCAP];
    __shared__ int32_t s_base[THREADS + 1];
    __shared__ int32_t s_j[THREADS];
    __shared__ uint8_t s_row[THREADS];
    __shared__ uint32_t s_em[THREADS];
    __shared__ int s_wave[WAVES];

    const int tid = threadIdx.x;
    const int64_t row0 = tile * ROWS;
    const int rows = static_cast<int>(min<int64_t>(ROWS, A.h - row0));
    const bool small = A.h <= 1024;
    const uint64_t n64 = static_cast<uint64_t>(A.n);
    const uint64_t modulus = small ? n64 : A.range;
    const bool windowed = !small && static_cast<uint64_t>(A.range) < n64;
    const bool retry = small && A.exclude;
    const double mod_d = static_cast<double>(static_cast<uint32_t>(modulus));
    const double inv_mod = 1.0 / mod_d;
    const uint64_t prod_limit = exact_product_limit(inv_mod);
    const uint64_t t1 = (static_cast<uint64_t>(A.epoch) + 1u) % modulus;
    int len = 0;
    if (tid < rows) {
        const int64_t row = row0 + tid;
        int64_t rk = 0;
        s_pos[tid] = ld2(A.pos, row);
        if (windowed) rk = static_cast<uint32_t>(A.ranks[row]);
        uint64_t first = 0;
        if (windowed) {
            const uint64_t half = A.range / 2;
            if (static_cast<uint64_t>(rk) > half) first = static_cast<uint64_t>(rk) - half;
            if (first > n64 - A.range) first = n64 - A.range;
        }
        const int32_t seg = A.offsets[row];
        len = A.offsets[row + 1] - seg;
        s_seg[tid] = seg;
        s_rowid[tid] = static_cast<int32_t>(row);
        s_first[tid] = static_cast<uint32_t>(first);
        s_active[tid] = A.block_of_row ? (A.block_of_row[row] == A.block_id) : 1;
    }
    if (tid < 64) {
        int incl = len;
        for (int off = 1; off < 64; off <<= 1) {
            const int up = __shfl_up(incl, off, 64);
            if (tid >= off) incl += up;
        }
        if (tid < ROWS) s_off[tid + 1] = incl;
        if (tid == 0) s_off[0] = 0;
    }
    const float scale = A.scale_ptr ? *A.scale_ptr : A.scale;
    __syncthreads();

    const int32_t e_begin = s_off[0], e_end = s_off[rows];
    const double now = static_cast<double>(A.epoch);
    const float fm = small ? 0.5f : 1.0f;
    const float s_neg = small ? 1.0f : scale;
    float2 acc = make_float2(0.0f, 0.0f);

    // Stream registers: the next chunk's clock/index/period words are
    // loaded while the current chunk's gathers are in flight.
    double pf_npos = 0.0, pf_per = 0.0, pf_nneg = 0.0, pf_nper = 1.0;
    int32_t pf_idx = 0, pf_e = 0, pf_r = 0;
    auto prefetch = [&](int32_t chunk) {
        const int32_t v = chunk + tid;
        if (v < e_end) {
            int lo = 0, hi = rows;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_off[mid] <= v) lo = mid;
                else hi = mid;
            }
            pf_r = lo;
            pf_e = s_seg[lo] + (v - s_off[lo]);
            pf_idx = A.indices[pf_e];
            if (A.sched) {
                // term count from the clock schedule: 1 + negatives, 0 when idle
                pf_npos = static_cast<double>(A.sched[pf_e]);
                return;
            }
            pf_npos = A.next_positive[pf_e];
            pf_per = A.periods[pf_e];
            if (A.negatives) {
                pf_nneg = A.next_negative[pf_e];
                pf_nper = A.negative_periods[pf_e];
            }
        }
    };
    prefetch(e_begin);

    for (int32_t base = e_begin; base < e_end; base += THREADS) {
        // ---- A: clocks and term counts
        const bool has_edge = base + tid < e_end;
        const int32_t e = pf_e;
        const int r_cur = pf_r;
        const double npos = pf_npos, per = pf_per, nn = pf_nneg, np = pf_nper;
        const int32_t jidx = pf_idx;
        if (base + THREADS < e_end) prefetch(base + THREADS);
        int terms = 0, count = 0, r = 0;
        float2 retry_sum = make_float2(0.0f, 0.0f);
        if (has_edge) {
            r = r_cur;
            if (s_active[r] && (A.sched ? npos != 0.0 : npos <= now)) {
                s_j[tid] = jidx;
                if (A.sched) {
                    count = static_cast<int32_t>(npos) - 1;
                } else {
                    A.next_positive[e] = npos + per;
                    if (A.negatives) {
                        count = static_cast<int32_t>((now - nn) / np);
                        A.next_negative[e] = nn + static_cast<double>(count) * np;
                    }
                }
                const int negs = count > 0 ? count : 0;
                if (retry && negs > 0) {
                    // Small layouts with exclusion retry rejected candidates
                    // sequentially; the accepted forces form one term.
                    NegWalk w;
                    w.init(static_cast<uint64_t>(e), A.epoch, negs, true, true, n64, A.range, 0);
                    const int64_t i = s_rowid[r];
                    const float2 xi = s_pos[r];
                    const int64_t seg = s_seg[r], seg_end = seg + (s_off[r + 1] - s_off[r]);
                    uint64_t rank;
                    while (w.next(rank)) {
                        const int64_t jj = static_cast<uint32_t>(A.order[rank]);
                        if (excluded(A.indices, seg, seg_end, i, jj, A.shared != 0)) continue;
                        w.accept();
                        const float2 y = A.table ? ld2(A.table, rank) : ld2(A.tgt, jj);
                        const float dx = xi.x - y.x, dy = xi.y - y.y;
                        const float c = repel_fast(dx * dx + dy * dy, A.a, A.b, A.gamma, 1.0f);
                        retry_sum.x = retry_sum.x + c * dx;
                        retry_sum.y = retry_sum.y + c * dy;
                    }
                    terms = 2;
                } else {
                    terms = 1 + negs;
                }
                s_em[tid] = static_cast<uint32_t>(static_cast<uint32_t>(e) % static_cast<uint32_t>(modulus));
            }
        }
        s_row[tid] = static_cast<uint8_t>(r);
        int total;
        const int my_base = block_exclusive_scan<WAVES>(terms, tid, s_wave, total);
        s_base[tid] = my_base;
        if (tid == THREADS - 1) s_base[THREADS] = total;

        // The chunk's terms are processed in windows of TCAP (almost always
        // one): B computes a window's terms term-parallel, C adds them to
        // their rows in term order. The per-row sum is therefore the
        // sequential per-term sum whatever the tile shape or term count.
        for (int w0 = 0; w0 < total; w0 += TCAP) {
            const int wend = min(total, w0 + TCAP);
            for (int k = max(my_base, w0); k < min(my_base + terms, wend); ++k)
                s_owner[k - w0] = static_cast<uint8_t>(tid);
            if (retry && terms == 2 && my_base + 1 >= w0 && my_base + 1 < wend)
                s_terms[my_base + 1 - w0] = retry_sum;
            __syncthreads();

            // ---- B: force terms (two per thread in flight)
            for (int item0 = w0 + tid; item0 < wend; item0 += 2 * THREADS) {
                float2 y[2], xi[2];
                int kk[2];
                bool live[2], keep[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int item = item0 + u * THREADS;
                    live[u] = item < wend;
                    keep[u] = live[u];
                    kk[u] = 0;
                    y[u] = make_float2(0.0f, 0.0f);
                    xi[u] = make_float2(0.0f, 0.0f);
                    if (!live[u]) continue;
                    const int owner = s_owner[item - w0];
                    const int k = item - s_base[owner];
                    const int rr = s_row[owner];
                    kk[u] = k;
                    xi[u] = s_pos[rr];
                    if (k == 0) {
                        y[u] = ld2(A.tgt, static_cast<uint32_t>(s_j[owner]));
                    } else if (retry) {
                        live[u] = false;  // combined retry term already stored
                    } else {
                        const uint64_t sel =
                            negative_offset(s_em[owner], static_cast<uint32_t>(t1), k, modulus,
                                            mod_d, inv_mod, prod_limit);
                        const uint64_t rank = static_cast<uint64_t>(s_first[rr]) + sel;
                        if (A.exclude) {
                            const int64_t jj = static_cast<uint32_t>(A.order[rank]);
                            const int64_t seg = s_seg[rr];
                            keep[u] = !excluded(A.indices, seg, seg + (s_off[rr + 1] - s_off[rr]),
                                                s_rowid[rr], jj, A.shared != 0);
                            y[u] = A.table ? ld2(A.table, rank) : ld2(A.tgt, jj);
                        } else {
                            y[u] = A.table ? ld2(A.table, rank)
                                           : ld2(A.tgt, static_cast<uint32_t>(A.order[rank]));
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    if (!live[u]) continue;
                    float2 term = make_float2(0.0f, 0.0f);
                    const float dx = xi[u].x - y[u].x, dy = xi[u].y - y[u].y;
                    const float q = dx * dx + dy * dy;
                    if (kk[u] == 0) {
                        if (q > 0.0f) {
                            const float c = attract_fast(q, A.a, A.b, fm);
                            term = make_float2(c * dx, c * dy);
                        }
                    } else if (keep[u]) {
                        const float c = repel_fast(q, A.a, A.b, A.gamma, s_neg);
                        term = make_float2(c * dx, c * dy);
                    }
                    s_terms[item0 + u * THREADS - w0] = term;
                }
            }
            __syncthreads();

            // ---- C: per-row sums in edge (and term) order
            if (tid < rows) {
                const int32_t lo = max(s_off[tid], base) - base;
                const int32_t hi = min(s_off[tid + 1], base + THREADS) - base;
                if (lo < hi) {
                    const int t_lo = max(s_base[lo], w0), t_hi = min(s_base[hi], wend);
                    for (int k = t_lo; k < t_hi; ++k) {
                        const float2 c = s_terms[k - w0];
                        acc.x = acc.x + c.x;
                        acc.y = acc.y + c.y;
                    }
                }
            }
            __syncthreads();
        }
    }

    if (tid >= rows || !s_active[tid]) return;
    const int64_t i = s_rowid[tid];
    if (A.fused) {
        const float2 x = s_pos[tid];
        float2 mm = ld2(A.m, i), vv = ld2(A.v, i);
        float2 out = x;
        if (acc.x != 0.0f) {
            mm.x = 0.9f * mm.x + (1.0f - 0.9f) * acc.x;
            vv.x = 0.99f * vv.x + ((1.0f - 0.99f) * acc.x) * acc.x;
            out.x = x.x + (A.alpha * (mm.x / A.c1)) / (sqrtf(vv.x / A.c2) + 0.0001f);
        }
        if (acc.y != 0.0f) {
            mm.y = 0.9f * mm.y + (1.0f - 0.9f) * acc.y;
            vv.y = 0.99f * vv.y + ((1.0f - 0.99f) * acc.y) * acc.y;
            out.y = x.y + (A.alpha * (mm.y / A.c1)) / (sqrtf(vv.y / A.c2) + 0.0001f);
        }
        reinterpret_cast<float2 *>(A.m)[i] = mm;
        reinterpret_cast<float2 *>(A.v)[i] = vv;
        reinterpret_cast<float2 *>(A.pos_out)[i] = out;
    } else {
        float2 u = ld2(A.updates, i);
        if (acc.x != 0.0f || acc.y != 0.0f) {
            u.x = u.x + A.multiplier * acc.x;
            u.y = u.y + A.multiplier * acc.y;
            reinterpret_cast<float2 *>(A.updates)[i] = u;
        }
    }
}

__global__ __launch_bounds__(kThreads) void edge_sweep_kernel(SweepArgs A) {
    sweep_tile<kTileRows, kThreads>(A, blockIdx.x);
}

// Many independent layouts in one launch: workgroup b runs tile
// b - tile0[s] of the segment s whose tile range holds b (segments are in
// ascending tile0 order). Each segment carries its own arguments, so its
// rows are computed exactly as a launch of that segment alone would.
template <int ROWS, int THREADS>
__global__ __launch_bounds__(THREADS) void segment_sweep_kernel(const SweepSegment *segs,
                                                                int32_t n_segs) {
    const int64_t b = blockIdx.x;
    int lo = 0, hi = n_segs;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (segs[mid].tile0 <= b) lo = mid;
        else hi = mid;
    }
    const SweepArgs A = segs[lo].args;
    const int64_t tile = b - segs[lo].tile0;
    if (tile * ROWS >= A.h) return;
    sweep_tile<ROWS, THREADS>(A, tile);
}

// Batched layouts in one persistent launch: every epoch of every segment.
// Epoch t's segments are segs[t * n_segs, (t + 1) * n_segs). Its tiles are
// split into 8 contiguous ranges, one per XCD; workgroups take chunks of
// kPersistChunk tiles from their XCD's range through that range's counter
// (ctr[t * 8 + range], zero at launch), then from the other ranges, so slow
// tiles (hub rows) do not hold the epoch back. Each tile runs as the
// per-epoch launch would (sweep_tile with the same arguments); a grid
// barrier separates the epochs. Launched cooperatively, so every workgroup
// is resident and the barrier cannot wait on a workgroup that never runs.
// Every counter only grows, so every workgroup leaves the chunk loops.
constexpr int kPersistChunk = 4;

template <int ROWS, int THREADS>
__global__ __launch_bounds__(THREADS, THREADS == 64 ? 5 : 4) void persistent_segment_kernel(
    const SweepSegment *segs,
                                                                     int32_t n_segs,
                                                                     int64_t total_tiles,
                                                                     int32_t epochs,
                                                                     int32_t *ctr) {
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    __shared__ int32_t s_take;
    const int xcd = static_cast<int>(blockIdx.x % 8);
    for (int32_t t = 0; t < epochs; ++t) {
        const SweepSegment *es = segs + static_cast<int64_t>(t) * n_segs;
        int seg = 0;
        int64_t seg_lo = 0, seg_hi = -1;   // tiles of the cached segment
        for (int r = 0; r < 8; ++r) {
            const int range = (xcd + r) & 7;
            const int64_t lo_t = total_tiles * range / 8, hi_t = total_tiles * (range + 1) / 8;
            for (;;) {
                if (threadIdx.x == 0) s_take = atomicAdd(&ctr[t * 8 + range], kPersistChunk);
                __syncthreads();
                const int64_t c0 = lo_t + s_take;
                __syncthreads();
                if (c0 >= hi_t) break;
                const int64_t c1 = min<int64_t>(c0 + kPersistChunk, hi_t);
                for (int64_t b = c0; b < c1; ++b) {
                    if (b < seg_lo || b >= seg_hi) {
                        int lo = 0, hi = n_segs;
                        while (hi - lo > 1) {
                            const int mid = (lo + hi) >> 1;
                            if (es[mid].tile0 <= b) lo = mid;
                            else hi = mid;
                        }
                        seg = lo;
                        seg_lo = es[lo].tile0;
                        seg_hi = lo + 1 < n_segs ? es[lo + 1].tile0 : total_tiles;
                    }
                    const SweepArgs A = es[seg].args;
                    const int64_t tile = b - seg_lo;
                    if (tile * ROWS < A.h) sweep_tile<ROWS, THREADS>(A, tile);
                    __syncthreads();   // the next tile reuses the LDS tables
                }
            }
        }
        grid.sync();
    }
}

// Transform layouts: workgroup b runs every epoch of tile b - tile0[s] of
// the request s whose tile range holds b. Epoch t's arguments are those of
// the per-epoch launch: the schedule, the Jacobi buffers t & 1 -> (t + 1) & 1
// and the negative tables of the epoch's direction (layouts of at most 1024
// rows: identity order at epoch 0, direction t - 1 after; larger layouts:
// direction t, with each row's window rank the lower bound of its own
// projection among the tail's sorted projections, as the per-epoch
// projection and lower-bound launches compute it). A row's positions, Adam
// moments, clocks and rank are read and written by the same lanes in every
// epoch, so no cross-workgroup ordering is needed.
constexpr int kTransformRows = 4;
constexpr int kTransformThreads = 64;

template <int ROWS, int THREADS>
__global__ __launch_bounds__(THREADS) void transform_layout_kernel(const TransformSegment *segs,
                                                                   int32_t n_segs,
                                                                   TransformTables T) {
    const int64_t b = blockIdx.x;
    int lo = 0, hi = n_segs;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (segs[mid].tile0 <= b) lo = mid;
        else hi = mid;
    }
    const TransformSegment &S = segs[lo];
    SweepArgs A = S.args;
    const int64_t tile = b - S.tile0;
    if (tile * ROWS >= A.h) return;
    float *buf[2] = {const_cast<float *>(A.pos), A.pos_out};
    const bool small = A.h <= 1024;
    const int64_t n = T.n;
    const int tid = threadIdx.x;
    const int64_t row = tile * ROWS + tid;
    const bool own = tid < ROWS && row < A.h;
    for (int32_t t = 0; t < T.epochs; ++t) {
        A.pos = buf[t & 1];
        A.pos_out = buf[(t + 1) & 1];
        A.epoch = static_cast<uint32_t>(t);
        A.alpha = T.alpha[t];
        A.c1 = T.c1[t];
        A.c2 = T.c2[t];
        if (small) {
            A.table = t == 0 ? T.tail : T.table + (t - 1) * n * 2;
            A.order = t == 0 ? T.iota : T.order + (t - 1) * n;
            A.ranks = nullptr;
        } else {
            const float *sorted = T.sorted + t * n;
            A.table = T.table + t * n * 2;
            A.order = T.order + t * n;
            A.ranks = S.ranks;
            if (own) {
                const float2 p = ld2(A.pos, row);
                float sum = 0.0f;
                sum += p.x * T.dirs[2 * t];
                sum += p.y * T.dirs[2 * t + 1];
                if (!isfinite(sum))
                    atomicMin(S.status, (static_cast<unsigned long long>(t) << 32) |
                                            static_cast<unsigned long long>(row));
                int64_t l = 0, h = n;
                while (l < h) {
                    const int64_t mid = l + (h - l) / 2;
                    if (sorted[mid] < sum) l = mid + 1;
                    else h = mid;
                }
                S.ranks[row] = static_cast<int32_t>(l < n - 1 ? l : n - 1);
            }
        }
        sweep_tile<ROWS, THREADS>(A, tile);
        __syncthreads();
    }
}

// ------------------------------------------------- edge-balanced unit sweep
// Adam layouts with H > 1024 and d = 2. One wave owns kUnitEdges consecutive
// CSR edges (a unit), so a hub row with thousands of edges is spread over
// many waves instead of serialising one tile, and every stream is read
// contiguously. Per unit:
//   1. the unit's row range and the firing shadow of its edges;
//   2. the records of the firing edges and the unit rows' positions/windows;
//   3. force terms laid out term-parallel in edge order, all gathers of a
//      group of 64*kUnitGroups terms in flight together, and per-row sums
//      by a segmented wave scan into LDS accumulators.
// A row wholly inside the unit stores its gradient; a row crossing a unit
// boundary stores one fragment per unit, which the Adam pass adds in unit
// order. Deterministic; no atomics.
//
// Edge record (32 B): next_positive, next_negative, period, and a word with
// the column index (low 32 bits) and e mod M (high 32 bits). The negative
// period is recomputed as (period / rate) * 1.5, the reference's formula; a
// plan takes this path only when every stored negative period equals it.
// The firing shadow holds min(max(ceil(next_positive), 0), 65535): for an
// integer epoch t, next_positive <= t exactly when the shadow is <= t. The
// id-order sweep reads it to load only the records of firing edges; the
// rank-ordered sweep reads every record of its rows (a row's records are
// contiguous) and tests next_positive <= t itself, which saves the
// dependent shadow load. With a clock schedule (unit_schedule_kernel, the
// default) both orders read one term-count byte and the CSR column per edge
// instead, and neither the records nor the shadow.
// Per-term cost is not what bounds these sweeps: each unit waits on a chain
// of dependent loads (rows, edge words, gathers), so occupancy (registers
// and LDS) and the number of round trips set the rate.
constexpr int kUnitEPL = 4;
constexpr int kUnitEdges = 64 * kUnitEPL;
constexpr int kUnitTermCap = kUnitEdges * 4;
constexpr int kUnitGroups = 4;
constexpr int kUnitNarrowRows = 64;

__device__ inline uint16_t fire_shadow(double next_positive) {
    const double c = ceil(next_positive);
    if (!(c < 65535.0)) return 65535;   // also NaN: never fires
    return c > 0.0 ? static_cast<uint16_t>(c) : 0;
}

// One step of a wave-wide segmented scan through DPP: the value of the lane
// CTRL selects (row_shr:n within 16-lane rows, row_bcast:15 / row_bcast:31
// across them) for the rows ROWMASK enables, else 0.
template <int CTRL, int ROWMASK>
__device__ inline void seg_scan_step(float &sx, float &sy, int rid) {
    const float ux = __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(sx), CTRL, ROWMASK, 0xf, false));
    const float uy = __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(sy), CTRL, ROWMASK, 0xf, false));
    const int ur = __builtin_amdgcn_update_dpp(0, rid, CTRL, ROWMASK, 0xf, false);
    if (ur == rid) {
        sx = sx + ux;
        sy = sy + uy;
    }
}

// Adds the segmented (by row) inclusive sums of one 64-term group into the
// row accumulators; `row` is -1 for lanes without a term. Rows are
// non-decreasing over the lanes that hold terms, so a partial sum ending at
// a lower lane belongs to this lane's row exactly when that lane's row is
// this lane's row. Lanes without a term carry zeros under row id 0; a
// source lane outside the enabled rows reads as (0, 0, id 0).
__device__ inline void segmented_accumulate(float2 term, int row, int lane, float2 *s_acc) {
    float sx = term.x, sy = term.y;
    const int rid = row + 1;
    seg_scan_step<0x111, 0xf>(sx, sy, rid);   // row_shr:1
    seg_scan_step<0x112, 0xf>(sx, sy, rid);   // row_shr:2
    seg_scan_step<0x114, 0xf>(sx, sy, rid);   // row_shr:4
    seg_scan_step<0x118, 0xf>(sx, sy, rid);   // row_shr:8
    seg_scan_step<0x142, 0xa>(sx, sy, rid);   // row_bcast:15 into rows 1 and 3
    seg_scan_step<0x143, 0xc>(sx, sy, rid);   // row_bcast:31 into rows 2 and 3
    const int next = __shfl_down(row, 1, 64);
    if (row >= 0 && (lane == 63 || next != row)) {
        float2 a = s_acc[row];
        a.x = a.x + sx;
        a.y = a.y + sy;
        s_acc[row] = a;
    }
}

// RANKED: units run over the rows in the epoch's projection order (rank r
// covers sequence slots [pref[r], pref[r+1])), so the heads a wave, and an
// XCD, sweep together have neighbouring ranks and their negative windows
// overlap in L2. Requires every row to have an edge.
// RC: capacity of the unit row table; the plan picks 64 when no unit can
// hold more rows, which keeps the workgroup at about 7 KB of LDS (6 waves
// per SIMD) instead of about 14 KB (3 waves).
// FMT (rank-ordered only for 1): how a firing edge is found and its terms
// counted.
//   0: 32-byte records (above), clocks advanced here;
//   1: 24-byte records, the period rebuilt as per_num / samples and e mod M
//      from the edge index, clocks advanced here;
//   2: the epoch's slab of the clock schedule (unit_schedule_kernel) gives
//      the term count of each edge and the CSR columns give its tail; the
//      records are not touched (the schedule kernel has advanced them).
// TRACE: also records every drawn negative (candidate id and whether it
// repels) per CSR edge, with the same selection code (tests).
template <bool RANKED, int RC, int FMT, bool TRACE>
__global__ __launch_bounds__(64) void unit_sweep_kernel(UnitSweepArgs U) {
    constexpr bool COMPACT = FMT == 1, SCHED = FMT == 2;
    static_assert(RANKED || FMT != 1, "compact records are read by rank only");
    const SweepArgs &A = U.s;
    constexpr int EPL = kUnitEPL, UE = kUnitEdges, TCAP = kUnitTermCap, G = kUnitGroups;
    __shared__ uint16_t s_off[RC + 1];  // unit-local first edge of each unit row
    __shared__ int32_t s_seg[RC + 1];   // CSR start of each unit row (+ end of the last)
    __shared__ int32_t s_rowid[RC];
    __shared__ float2 s_pos[RC];
    __shared__ uint32_t s_first[RANKED ? 1 : RC];   // window start of each unit row
    __shared__ float2 s_acc[RC];
    __shared__ uint8_t s_owner[TCAP];   // edge slot of each term
    __shared__ uint16_t s_base[UE + 1]; // first term of each edge slot (read when total <= TCAP)
    __shared__ int32_t s_j[UE];
    __shared__ uint8_t s_row[UE];
    __shared__ uint32_t s_em[SCHED ? 1 : UE];   // e mod M (SCHED: recomputed per term)
    __shared__ int32_t s_pst[RANKED ? RC + 1 : 1];  // sequence start of each unit row
    __shared__ int32_t s_end[RANKED ? RC + 1 : 1];  // CSR end of each unit row
    static_assert(UE <= 256 && RC <= UE, "edge slots and unit rows are indexed by uint8");
    static_assert(TCAP < 65536, "term bases are 16-bit");

    const int lane = threadIdx.x;
    // Workgroups are dealt to the 8 XCDs in turn; workgroup b takes unit
    // (b mod 8) * xcd_chunk + b / 8, so each XCD (and its L2) sweeps one
    // contiguous eighth of the rows.
    const int64_t u = (blockIdx.x % 8) * U.xcd_chunk + blockIdx.x / 8;
    if (u >= U.units) return;
    const int64_t e0 = u * UE;
    const int ue = static_cast<int>(min<int64_t>(UE, U.n_edges - e0));
    const uint32_t epoch = A.epoch;
    const uint64_t n64 = static_cast<uint64_t>(A.n);
    const uint64_t modulus = A.range;
    const bool windowed = modulus < n64;
    const double mod_d = static_cast<double>(static_cast<uint32_t>(modulus));
    const uint64_t prod_limit = exact_product_limit(U.inv_mod);
    const uint32_t t1 = static_cast<uint32_t>(U.t1);
    // 1. unit rows and firing shadow
    // first and last row (ranked: rank) of the unit
    const int32_t k0 = U.unit_k[2 * u], k1 = U.unit_k[2 * u + 1];
    const int nr = k1 - k0 + 1;
    uint32_t fire[EPL];
    int64_t eidx[EPL];
    double r_npos[EPL], r_nn[EPL], r_per[EPL];
    uint64_t r_w[EPL];
    if constexpr (!RANKED && SCHED) {
        // the unit's edges are contiguous: term counts and columns at once
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            const int v = k * 64 + lane;
            eidx[k] = e0 + v;
            fire[k] = 0;   // term count (0: the edge does not fire)
            r_w[k] = 0;
            if (v < ue) {
                fire[k] = U.sched[e0 + v];
                r_w[k] = static_cast<uint32_t>(A.indices[e0 + v]);
            }
        }
    } else if constexpr (!RANKED) {
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            const int v = k * 64 + lane;
            eidx[k] = e0 + v;
            fire[k] = v < ue ? U.fire[e0 + v] : 0xffffffffu;
        }
        // 2. firing records, then the row data
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            r_npos[k] = 0.0; r_nn[k] = 0.0; r_per[k] = 0.0; r_w[k] = 0;
            if (fire[k] <= epoch) {
                const double4 rec = reinterpret_cast<const double4 *>(U.rec)[eidx[k]];
                r_npos[k] = rec.x;
                r_nn[k] = rec.y;
                r_per[k] = rec.z;
                r_w[k] = static_cast<uint64_t>(__double_as_longlong(rec.w));
            }
        }
    }
    for (int lr = lane; lr < nr; lr += 64) {
        const int32_t k = k0 + lr;
        int32_t row, seg, pst;
        if constexpr (RANKED) {
            // everything by rank: no load waits on the row id
            row = A.order[k];
            seg = U.rank_seg[k];
            pst = U.pref[k];
            s_pst[lr] = pst;
            s_end[lr] = seg + (U.pref[k + 1] - pst);
            s_pos[lr] = ld2(A.table, k);   // the position of order[k]
        } else {
            row = U.nz_rows ? U.nz_rows[k] : k;
            seg = U.nz_off[k];
            pst = seg;
            s_pos[lr] = ld2(A.pos, row);
        }
        s_seg[lr] = seg;
        s_off[lr] = static_cast<uint16_t>(max<int64_t>(pst, e0) - e0);
        s_rowid[lr] = row;
        uint64_t first = 0;
        if (windowed) {
            const uint64_t rk = RANKED ? static_cast<uint64_t>(k)
                                       : static_cast<uint64_t>(static_cast<uint32_t>(A.ranks[row]));
            const uint64_t half = modulus / 2;
            if (rk > half) first = rk - half;
            if (first > n64 - modulus) first = n64 - modulus;
        }
        if constexpr (!RANKED) s_first[lr] = static_cast<uint32_t>(first);
        s_acc[lr] = make_float2(0.0f, 0.0f);
    }
    if (lane == 0) {
        s_off[nr] = static_cast<uint16_t>(ue);
        if constexpr (!RANKED) s_seg[nr] = U.nz_off[k1 + 1];
    }
    // window start of unit row rr (ranked: from its rank k0 + rr)
    auto row_first = [&](int rr) -> uint64_t {
        if constexpr (RANKED) {
            if (!windowed) return 0;
            const uint64_t rk = static_cast<uint64_t>(k0 + rr), half = modulus / 2;
            const uint64_t first = rk > half ? rk - half : 0;
            return first > n64 - modulus ? n64 - modulus : first;
        } else {
            return s_first[rr];
        }
    };
    const float scale = A.scale_ptr ? *A.scale_ptr : A.scale;
    const double now = static_cast<double>(epoch);
    const bool repel_on = A.gamma > 0.0f;
    const float k_att = -2.0f * A.a * A.b;
    const float k_rep = scale * 2.0f * A.gamma * A.b;
    const float inv_gamma = rcpf(A.gamma);
    const float g16 = A.gamma < 16.0f ? A.gamma : 16.0f;
    __syncthreads();
    if constexpr (RANKED) {
        // slots to CSR edges through the row table, then the records
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            const int slot = k * 64 + lane;
            eidx[k] = 0;
            fire[k] = 0xffffffffu;
            if (slot < ue) {
                int lo = 0, hi = nr;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (s_off[mid] <= slot) lo = mid;
                    else hi = mid;
                }
                s_row[slot] = static_cast<uint8_t>(lo);   // read back by this lane only
                eidx[k] = s_seg[lo] + (e0 + slot - s_pst[lo]);
            }
        }
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            r_npos[k] = 0.0; r_nn[k] = 0.0; r_per[k] = 0.0; r_w[k] = 0;
            if constexpr (SCHED) {
                // fire[k] holds the edge's term count (0: it does not fire)
                fire[k] = 0;
                if (k * 64 + lane < ue) {
                    fire[k] = U.sched[eidx[k]];
                    r_w[k] = static_cast<uint32_t>(A.indices[eidx[k]]);
                }
                continue;
            }
            if (k * 64 + lane < ue) {
                if constexpr (COMPACT) {
                    const double *rec = U.rec + 3 * eidx[k];
                    r_npos[k] = rec[0];
                    r_nn[k] = rec[1];
                    r_w[k] = static_cast<uint64_t>(__double_as_longlong(rec[2]));
                } else {
                    const double4 rec = reinterpret_cast<const double4 *>(U.rec)[eidx[k]];
                    r_npos[k] = rec.x;
                    r_nn[k] = rec.y;
                    r_per[k] = rec.z;
                    r_w[k] = static_cast<uint64_t>(__double_as_longlong(rec.w));
                }
                fire[k] = r_npos[k] <= now ? 0u : 0xffffffffu;
            }
        }
    }

    // clocks (bit-exact f64, reference order of operations) and term counts
    int running = 0;
    int t_cnt[EPL], t_base[EPL];
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
        const int slot = k * 64 + lane;
        int t = 0, r = 0;
        if (slot < ue) {
            if constexpr (RANKED) {
                r = s_row[slot];
            } else {
                int lo = 0, hi = nr;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (s_off[mid] <= slot) lo = mid;
                    else hi = mid;
                }
                r = lo;
            }
            if constexpr (SCHED) {
                if (fire[k] != 0) {
                    const int64_t e = eidx[k];
                    if constexpr (TRACE) {
                        const int negs = static_cast<int>(fire[k]) - 1;
                        U.tr_count[e] = negs;
                        if (negs > U.tr_cap) atomicOr(U.tr_overflow, 1);
                    }
                    t = static_cast<int>(fire[k]);
                    s_j[slot] = static_cast<int32_t>(static_cast<uint32_t>(r_w[k]));
                }
            } else if (fire[k] <= epoch) {
                const int64_t e = eidx[k];
                double per = r_per[k];
                if constexpr (COMPACT)
                    per = U.per_num / static_cast<double>(
                                          __uint_as_float(static_cast<uint32_t>(r_w[k] >> 32)));
                const double nn = r_nn[k];
                const double npos2 = r_npos[k] + per;
                int32_t count = 0;
                double nn2 = nn;
                if (A.negatives) {
                    const double np = (per / U.rate) * 1.5;
                    count = static_cast<int32_t>((now - nn) / np);
                    nn2 = nn + static_cast<double>(count) * np;
                }
                if constexpr (COMPACT) {
                    U.rec[3 * e] = npos2;
                    U.rec[3 * e + 1] = nn2;
                } else {
                    reinterpret_cast<double2 *>(U.rec)[2 * e] = make_double2(npos2, nn2);
                }
                if constexpr (!RANKED) U.fire[e] = fire_shadow(npos2);
                const int negs = count > 0 ? count : 0;
                if constexpr (TRACE) {
                    U.tr_count[e] = negs;
                    if (negs > U.tr_cap) atomicOr(U.tr_overflow, 1);
                }
                t = 1 + negs;
                // e < 2^31: exact in f64
                const uint32_t em =
                    COMPACT ? mod_f64(static_cast<double>(static_cast<uint32_t>(e)), mod_d, U.inv_mod)
                            : static_cast<uint32_t>(r_w[k] >> 32);
                s_j[slot] = static_cast<int32_t>(static_cast<uint32_t>(r_w[k]));
                s_em[slot] = em;
            }
        }
        s_row[slot] = static_cast<uint8_t>(r);
        const int incl = wave_inclusive_sum(t);
        t_cnt[k] = t;
        t_base[k] = running + incl - t;
        s_base[slot] = static_cast<uint16_t>(t_base[k]);
        running += __builtin_amdgcn_readlane(incl, 63);
    }
    const int total = running;
    if (lane == 0) s_base[UE] = static_cast<uint16_t>(total);
    const bool term_parallel = total <= TCAP;
    if (term_parallel) {
#pragma unroll
        for (int k = 0; k < EPL; ++k)
            for (int q = 0; q < t_cnt[k]; ++q)
                s_owner[t_base[k] + q] = static_cast<uint8_t>(k * 64 + lane);
    }
    if constexpr (TRACE) {
        if (lane == 0) atomicAdd(U.tr_paths + (term_parallel ? 0 : 1), 1);
    }
    __syncthreads();

    if (term_parallel) {
        // 3. all terms of a pass gathered together, then computed and summed
        for (int p0 = 0; p0 < total; p0 += 64 * G) {
            float2 y[G];
            bool keep[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int item = p0 + g * 64 + lane;
                y[g] = make_float2(0.0f, 0.0f);
                keep[g] = true;
                if (item >= total) continue;
                const int owner = s_owner[item];
                const int kq = item - s_base[owner];
                if (kq == 0) {
                    y[g] = ld2(A.tgt, static_cast<uint32_t>(s_j[owner]));
                    continue;
                }
                const int rr = s_row[owner];
                uint32_t em;
                if constexpr (SCHED) {
                    // e < 2^31: exact in f64
                    const int64_t e = RANKED ? s_seg[rr] + (e0 + owner - s_pst[rr]) : e0 + owner;
                    em = mod_f64(static_cast<double>(static_cast<uint32_t>(e)), mod_d, U.inv_mod);
                } else {
                    em = s_em[owner];
                }
                const uint64_t sel =
                    negative_offset(em, t1, kq, modulus, mod_d, U.inv_mod, prod_limit);
                const uint64_t rank = row_first(rr) + sel;
                if (A.exclude) {
                    const int64_t jj = static_cast<uint32_t>(A.order[rank]);
                    keep[g] = !excluded(A.indices, s_seg[rr], RANKED ? s_end[rr] : s_seg[rr + 1],
                                        s_rowid[rr], jj, A.shared != 0);
                    y[g] = A.table ? ld2(A.table, rank) : ld2(A.tgt, jj);
                } else {
                    y[g] = A.table ? ld2(A.table, rank)
                                   : ld2(A.tgt, static_cast<uint32_t>(A.order[rank]));
                }
                if constexpr (TRACE) {
                    const int64_t e = RANKED ? s_seg[rr] + (e0 + owner - s_pst[rr]) : e0 + owner;
                    if (kq <= U.tr_cap) {
                        const int64_t at = e * U.tr_cap + (kq - 1);
                        U.tr_cand[at] = A.order[rank];
                        U.tr_kept[at] = keep[g] ? 1 : 0;
                    }
                }
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (p0 + g * 64 >= total) break;   // wave-uniform
                const int item = p0 + g * 64 + lane;
                float2 term = make_float2(0.0f, 0.0f);
                int row = -1;
                if (item < total) {
                    const int owner = s_owner[item];
                    const int kq = item - s_base[owner];
                    row = s_row[owner];
                    const float2 xi = s_pos[row];
                    const float dx = xi.x - y[g].x, dy = xi.y - y[g].y;
                    const float q = dx * dx + dy * dy;
                    const bool att = kq == 0;
                    if (q > 0.0f && (att || (keep[g] && repel_on))) {
                        const float c = unit_coef(q, att, A.a, A.b, k_att, k_rep, inv_gamma, g16);
                        term = make_float2(c * dx, c * dy);
                    }
                }
                segmented_accumulate(term, row, lane, s_acc);
            }
        }
    } else {
        // More terms than LDS holds: each lane sums its edges' terms in
        // order, then the per-edge sums are reduced per row.
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            const int slot = k * 64 + lane;
            float2 g = make_float2(0.0f, 0.0f);
            const int rr = slot < ue ? s_row[slot] : 0;
            if (t_cnt[k] > 0) {
                const float2 xi = s_pos[rr];
                const float2 y = ld2(A.tgt, static_cast<uint32_t>(s_j[slot]));
                const float dx = xi.x - y.x, dy = xi.y - y.y;
                const float q = dx * dx + dy * dy;
                if (q > 0.0f) {
                    const float c = attract_fast(q, A.a, A.b, 1.0f);
                    g = make_float2(c * dx, c * dy);
                }
                const uint64_t em =
                    SCHED ? mod_f64(static_cast<double>(static_cast<uint32_t>(eidx[k])), mod_d,
                                    U.inv_mod)
                          : s_em[slot];
                uint64_t sel = t_cnt[k] > 1 ? negative_offset(static_cast<uint32_t>(em), t1, 1,
                                                              modulus, mod_d, U.inv_mod, prod_limit)
                                            : 0u;
                for (int p = 0; p + 1 < t_cnt[k]; ++p) {
                    if (p > 0) {
                        sel += em;
                        if (sel >= modulus) sel -= modulus;
                    }
                    const uint64_t rank = row_first(rr) + sel;
                    float2 yn;
                    if constexpr (TRACE) {
                        if (p < U.tr_cap) {
                            const int64_t at = eidx[k] * U.tr_cap + p;
                            const int64_t jt = static_cast<uint32_t>(A.order[rank]);
                            U.tr_cand[at] = static_cast<int32_t>(jt);
                            U.tr_kept[at] =
                                A.exclude && excluded(A.indices, s_seg[rr],
                                                      RANKED ? s_end[rr] : s_seg[rr + 1],
                                                      s_rowid[rr], jt, A.shared != 0)
                                    ? 0 : 1;
                        }
                    }
                    if (A.exclude) {
                        const int64_t jj = static_cast<uint32_t>(A.order[rank]);
                        if (excluded(A.indices, s_seg[rr], RANKED ? s_end[rr] : s_seg[rr + 1],
                                     s_rowid[rr], jj, A.shared != 0))
                            continue;
                        yn = A.table ? ld2(A.table, rank) : ld2(A.tgt, jj);
                    } else {
                        yn = A.table ? ld2(A.table, rank)
                                     : ld2(A.tgt, static_cast<uint32_t>(A.order[rank]));
                    }
                    const float nx = xi.x - yn.x, ny = xi.y - yn.y;
                    const float c = repel_fast(nx * nx + ny * ny, A.a, A.b, A.gamma, scale);
                    g.x = g.x + c * nx;
                    g.y = g.y + c * ny;
                }
            }
            segmented_accumulate(g, slot < ue ? rr : -1, lane, s_acc);
        }
    }
    __syncthreads();
    for (int lr = lane; lr < nr; lr += 64) {
        const float2 g = s_acc[lr];
        int slot = -1;
        if constexpr (RANKED) {
            // a row begun in an earlier unit stores slot 2u; one that
            // continues into the next unit (and began here) stores 2u + 1,
            // and its flag and unit range for the Adam step
            if (lr == 0 && s_pst[0] < e0) {
                slot = static_cast<int>(2 * u);
            } else if (lr == nr - 1 && s_off[lr] + (s_end[lr] - s_seg[lr]) > ue) {
                slot = static_cast<int>(2 * u + 1);
                const int64_t last = (s_pst[lr] + (s_end[lr] - s_seg[lr]) - 1) / UE;
                reinterpret_cast<int2 *>(U.grad)[s_rowid[lr]] =
                    make_int2(static_cast<int>(u), static_cast<int>(last));
                U.split[s_rowid[lr]] = 1;   // cleared again by the Adam step
            }
        } else {
            if (lr == 0) slot = U.unit_frag[2 * u];
            if (lr == nr - 1 && slot < 0) slot = U.unit_frag[2 * u + 1];
        }
        if (slot >= 0) reinterpret_cast<float2 *>(U.frag)[slot] = g;
        else reinterpret_cast<float2 *>(U.grad)[s_rowid[lr]] = g;
    }
}

__global__ void unit_pack_kernel(const double *npos, const double *nneg, const double *periods,
                                 const int32_t *indices, int64_t E, uint64_t modulus, bool compact,
                                 double per_num, double *rec, uint16_t *fire) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= E) return;
    if (compact) {
        const float samples = static_cast<float>(per_num / periods[e]);
        const uint64_t w = (static_cast<uint64_t>(__float_as_uint(samples)) << 32) |
                           static_cast<uint32_t>(indices[e]);
        rec[3 * e] = npos[e];
        rec[3 * e + 1] = nneg[e];
        rec[3 * e + 2] = __longlong_as_double(static_cast<long long>(w));
    } else {
        const uint64_t em = static_cast<uint64_t>(e) % modulus;
        const uint64_t w = (em << 32) | static_cast<uint32_t>(indices[e]);
        reinterpret_cast<double4 *>(rec)[e] = make_double4(
            npos[e], nneg[e], periods[e], __longlong_as_double(static_cast<long long>(w)));
    }
    fire[e] = fire_shadow(npos[e]);
}

__global__ void unit_unpack_kernel(const double *rec, int64_t E, bool compact, double *npos,
                                   double *nneg) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int64_t at = compact ? 3 * e : 4 * e;
    npos[e] = rec[at];
    nneg[e] = rec[at + 1];
}

// Clock schedule of the unit sweep for epochs [t0, t0 + span):
// sched[(t - t0) * stride + e] = 1 + the negatives edge e draws when it
// fires in epoch t, 0 when it does not fire. Neither clock depends on the
// positions, so one pass runs both through the span with the f64
// operations of the sweep's own clock update, in the same order
// (modern_layout.rs:662-691): the edge fires when next_positive <= t, then
// next_positive += period; count = trunc((t - next_negative) / negative
// period), next_negative += count * negative period. The records are left
// at the state after the span's last epoch. A lane holds four consecutive
// edges and writes their bytes as one word per epoch (stride is a multiple
// of 4 and at least E rounded up to one). A term count above 255 is stored
// as 255 and ORs 4 into *flag.
//
// schedule_lanes runs the four edges e0..e0+3 of a lane through the span
// (edges past the end carry next_positive = inf and never fire) and returns
// whether a count overflowed.
__device__ inline bool schedule_lanes(double (&npos)[4], double (&nn)[4], const double (&per)[4],
                                      const double (&np)[4], bool negatives, int32_t t0,
                                      int32_t span, uint8_t *sched, int64_t stride, int64_t e0) {
    bool over = false;
    for (int32_t i = 0; i < span; ++i) {
        const double now = static_cast<double>(t0 + i);
        uint32_t word = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (npos[k] <= now) {
                npos[k] = npos[k] + per[k];
                int32_t count = 0;
                if (negatives) {
                    count = static_cast<int32_t>((now - nn[k]) / np[k]);
                    nn[k] = nn[k] + static_cast<double>(count) * np[k];
                }
                const int32_t terms = 1 + (count > 0 ? count : 0);
                over = over || terms > 255;
                word |= static_cast<uint32_t>(terms < 255 ? terms : 255) << (8 * k);
            }
        }
        *reinterpret_cast<uint32_t *>(sched + i * stride + e0) = word;
    }
    return over;
}

__global__ void unit_schedule_kernel(double *rec, int64_t E, bool compact, double per_num,
                                     double rate, bool negatives, int32_t t0, int32_t span,
                                     uint8_t *sched, int64_t stride, int32_t *flag) {
    const int64_t e0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
    if (e0 >= E) return;
    double npos[4], nn[4], per[4], np[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t e = e0 + k;
        npos[k] = INFINITY;   // past the end: never fires
        nn[k] = 0.0;
        per[k] = 1.0;
        if (e < E) {
            if (compact) {
                npos[k] = rec[3 * e];
                nn[k] = rec[3 * e + 1];
                const uint64_t w = static_cast<uint64_t>(__double_as_longlong(rec[3 * e + 2]));
                per[k] = per_num / static_cast<double>(__uint_as_float(static_cast<uint32_t>(w >> 32)));
            } else {
                npos[k] = rec[4 * e];
                nn[k] = rec[4 * e + 1];
                per[k] = rec[4 * e + 2];
            }
        }
        np[k] = negatives ? (per[k] / rate) * 1.5 : 1.0;
    }
    const bool over = schedule_lanes(npos, nn, per, np, negatives, t0, span, sched, stride, e0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t e = e0 + k;
        if (e >= E) break;
        const int64_t at = compact ? 3 * e : 4 * e;
        rec[at] = npos[k];
        rec[at + 1] = nn[k];
    }
    if (over) atomicOr(flag, 4);
}

// The same schedule over clocks kept as four f64 arrays (the tile sweep's
// form): the negative period is read, not rebuilt.
__global__ void edge_schedule_kernel(double *npos_a, double *nneg_a, const double *per_a,
                                     const double *nper_a, int64_t E, bool negatives, int32_t t0,
                                     int32_t span, uint8_t *sched, int64_t stride, int32_t *flag) {
    const int64_t e0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
    if (e0 >= E) return;
    double npos[4], nn[4], per[4], np[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t e = e0 + k;
        const bool in = e < E;
        npos[k] = in ? npos_a[e] : INFINITY;   // past the end: never fires
        nn[k] = in ? nneg_a[e] : 0.0;
        per[k] = in ? per_a[e] : 1.0;
        np[k] = in && negatives ? nper_a[e] : 1.0;
    }
    const bool over = schedule_lanes(npos, nn, per, np, negatives, t0, span, sched, stride, e0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t e = e0 + k;
        if (e >= E) break;
        npos_a[e] = npos[k];
        nneg_a[e] = nn[k];
    }
    if (over) atomicOr(flag, 4);
}

// Negative tables of a batch for epoch t: for row i of graph g (rows r0 ..
// r0 + h - 1), table[i] = pos[r0 + order_g,t[i - r0]], the position the
// graph's epoch-t order puts at that rank.
__global__ void batch_table_kernel(const float *pos, const int32_t *orders,
                                   const int64_t *row_start, const int64_t *order_start,
                                   const int32_t *graph_of_row, int64_t total_rows, int32_t t,
                                   float *table) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total_rows) return;
    const int32_t g = graph_of_row[i];
    const int64_t r0 = row_start[g], h = row_start[g + 1] - r0;
    const int64_t j = static_cast<uint32_t>(orders[order_start[g] + t * h + (i - r0)]);
    reinterpret_cast<float2 *>(table)[i] = ld2(pos, r0 + j);
}

// Lazy Adam over rows; rows split across units add their fragments in unit
// order first.
__global__ void unit_adam_kernel(int64_t h, const float *pos, float *pos_out, const float *grad,
                                 const int32_t *row_split, const int32_t *split_off,
                                 const float *frag, float *m, float *v, float alpha, float c1,
                                 float c2, float *grad_out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= h) return;
    const int32_t s = row_split[i];
    float2 g;
    if (s < 0) {
        g = ld2(grad, i);
    } else {
        g = make_float2(0.0f, 0.0f);
        for (int32_t q = split_off[s]; q < split_off[s + 1]; ++q) {
            const float2 f = ld2(frag, q);
            g.x = g.x + f.x;
            g.y = g.y + f.y;
        }
    }
    if (grad_out) reinterpret_cast<float2 *>(grad_out)[i] = g;
    const float2 x = ld2(pos, i);
    float2 mm = ld2(m, i), vv = ld2(v, i);
    float2 out = x;
    if (g.x != 0.0f) {
        mm.x = 0.9f * mm.x + (1.0f - 0.9f) * g.x;
        vv.x = 0.99f * vv.x + ((1.0f - 0.99f) * g.x) * g.x;
        out.x = x.x + (alpha * (mm.x / c1)) / (sqrtf(vv.x / c2) + 0.0001f);
    }
    if (g.y != 0.0f) {
        mm.y = 0.9f * mm.y + (1.0f - 0.9f) * g.y;
        vv.y = 0.99f * vv.y + ((1.0f - 0.99f) * g.y) * g.y;
        out.y = x.y + (alpha * (mm.y / c1)) / (sqrtf(vv.y / c2) + 0.0001f);
    }
    reinterpret_cast<float2 *>(m)[i] = mm;
    reinterpret_cast<float2 *>(v)[i] = vv;
    reinterpret_cast<float2 *>(pos_out)[i] = out;
}

// Lazy Adam for the rank-ordered sweep: a row inside one unit has its
// gradient in grad (split[i] = 0); otherwise (split[i] = 1, set by the
// sweep's first unit of the row and cleared here for the next epoch) grad
// holds the bits of its first and last unit u_a, u_b, and its fragments are
// slot 2u_a + 1 of the first unit and slot 2u of every later unit, added in
// unit order.
__global__ void unit_adam_ranked_kernel(int64_t h, const float *pos, float *pos_out,
                                        const float *grad, uint8_t *split, const float *frag,
                                        float *m, float *v, float alpha, float c1, float c2,
                                        float *grad_out, const int32_t *csr_off,
                                        float4 *rows_out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= h) return;
    float2 g = ld2(grad, i);
    if (split[i]) {
        split[i] = 0;
        const int64_t ua = __float_as_int(g.x), ub = __float_as_int(g.y);
        g = make_float2(0.0f, 0.0f);
        const float2 f0 = ld2(frag, 2 * ua + 1);
        g.x = g.x + f0.x;
        g.y = g.y + f0.y;
        for (int64_t w = ua + 1; w <= ub; ++w) {
            const float2 f = ld2(frag, 2 * w);
            g.x = g.x + f.x;
            g.y = g.y + f.y;
        }
    }
    if (grad_out) reinterpret_cast<float2 *>(grad_out)[i] = g;
    const float2 x = ld2(pos, i);
    float2 mm = ld2(m, i), vv = ld2(v, i);
    float2 out = x;
    if (g.x != 0.0f) {
        mm.x = 0.9f * mm.x + (1.0f - 0.9f) * g.x;
        vv.x = 0.99f * vv.x + ((1.0f - 0.99f) * g.x) * g.x;
        out.x = x.x + (alpha * (mm.x / c1)) / (sqrtf(vv.x / c2) + 0.0001f);
    }
    if (g.y != 0.0f) {
        mm.y = 0.9f * mm.y + (1.0f - 0.9f) * g.y;
        vv.y = 0.99f * vv.y + ((1.0f - 0.99f) * g.y) * g.y;
        out.y = x.y + (alpha * (mm.y / c1)) / (sqrtf(vv.y / c2) + 0.0001f);
    }
    reinterpret_cast<float2 *>(m)[i] = mm;
    reinterpret_cast<float2 *>(v)[i] = vv;
    reinterpret_cast<float2 *>(pos_out)[i] = out;
    if (rows_out) {
        const int32_t b = csr_off[i];
        rows_out[i] = make_float4(out.x, out.y, __int_as_float(b), __int_as_float(csr_off[i + 1] - b));
    }
}

// rows[i] = {x_i, y_i, CSR start, degree} (bits of the two integers), the
// record the projection sort's last pass gathers per row.
__global__ void pack_rows_kernel(const float *pos, const int32_t *csr_off, int64_t h,
                                 float4 *rows) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= h) return;
    const float2 x = ld2(pos, i);
    const int32_t b = csr_off[i];
    rows[i] = make_float4(x.x, x.y, __int_as_float(b), __int_as_float(csr_off[i + 1] - b));
}

// Exclusive prefix sum of the row degrees in projection order:
// pref[r] = sum over ranks below r of degree(order[r']), pref[n] = E.
// Pass 1 scans blocks of kPrefBlock ranks and writes each block's total;
// pass 2 adds to each rank the totals of the blocks before its own (every
// workgroup sums them itself: at most n / kPrefBlock values, in L2).
constexpr int kPrefThreads = 256, kPrefPer = 4, kPrefBlock = kPrefThreads * kPrefPer;

__device__ inline int32_t block_exclusive_scan(int32_t x, int32_t *s_warp, int32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int32_t incl = x;
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t up = __shfl_up(incl, o, 64);
        if (lane >= o) incl += up;
    }
    if (lane == 63) s_warp[wave] = incl;
    __syncthreads();
    int32_t before = 0;
    total = 0;
    for (int w = 0; w < kPrefThreads / 64; ++w) {
        if (w < wave) before += s_warp[w];
        total += s_warp[w];
    }
    __syncthreads();
    return before + incl - x;
}

__global__ __launch_bounds__(kPrefThreads) void rank_prefix_kernel(const int32_t *rank_deg,
                                                                   int64_t n, int32_t *pref,
                                                                   int32_t *block_sums) {
    __shared__ int32_t s_warp[kPrefThreads / 64];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kPrefBlock + threadIdx.x * kPrefPer;
    int32_t deg[kPrefPer], sum = 0;
#pragma unroll
    for (int q = 0; q < kPrefPer; ++q) {
        deg[q] = base + q < n ? rank_deg[base + q] : 0;
        sum += deg[q];
    }
    int32_t total;
    int32_t run = block_exclusive_scan(sum, s_warp, total);
#pragma unroll
    for (int q = 0; q < kPrefPer; ++q) {
        if (base + q < n) pref[base + q] = run;
        run += deg[q];
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Adds the block offsets and scatters each rank into the units whose first
// or last slot it holds (rank r covers slots [pref[r], pref[r] + deg)). A
// workgroup's ranks lie in one block (kPrefBlock is a multiple of
// kPrefThreads); it sums the totals of the blocks before it.
__global__ __launch_bounds__(kPrefThreads) void prefix_units_kernel(int32_t *pref, int64_t n,
                                                                    const int32_t *block_sums,
                                                                    const int32_t *rank_deg,
                                                                    int32_t total_edges,
                                                                    int32_t *unit_k) {
    static_assert(kPrefBlock % kPrefThreads == 0, "a workgroup's ranks lie in one block");
    __shared__ int32_t s_warp[kPrefThreads / 64];
    const int64_t r = static_cast<int64_t>(blockIdx.x) * kPrefThreads + threadIdx.x;
    const int64_t blk = static_cast<int64_t>(blockIdx.x) * kPrefThreads / kPrefBlock;
    int32_t part = 0;
    for (int64_t i = threadIdx.x; i < blk; i += kPrefThreads) part += block_sums[i];
    int32_t before;
    (void)block_exclusive_scan(part, s_warp, before);   // before = the workgroup's total
    if (r == 0) pref[n] = total_edges;
    if (r >= n) return;
    const int64_t b = pref[r] + before;
    pref[r] = static_cast<int32_t>(b);
    const int64_t en = b + rank_deg[r];
    if (en == b) return;
    const int64_t last_unit = (static_cast<int64_t>(total_edges) - 1) / kUnitEdges;
    for (int64_t u = b / kUnitEdges; u <= (en - 1) / kUnitEdges; ++u) {
        const int64_t first = u * kUnitEdges;
        const int64_t last = u == last_unit ? total_edges - 1 : first + kUnitEdges - 1;
        if (first >= b && first < en) unit_k[2 * u] = static_cast<int32_t>(r);
        if (last >= b && last < en) unit_k[2 * u + 1] = static_cast<int32_t>(r);
    }
}

// ------------------------------------------------ FP64 cross-entropy objective
// Fuzzy-set cross entropy of the embedding over the graph's positive edges:
//   sum_e -w_e ln(max(phi_e, eps)) - (1 - w_e) ln(max(1 - phi_e, eps)),
//   phi_e = 1 / (1 + a q_e^b), q_e = |x_i - y_j|^2, eps = 1e-4,
// every term and sum in FP64 (coordinates widened from f32). One lane per
// row sums its edges in order; workgroups reduce by a fixed tree and a
// second pass adds the workgroup partials in order, so the value is
// deterministic.
constexpr int kObjThreads = 256;

__device__ inline double cross_entropy_term(double q, double w, double a, double b) {
    const double p = q > 0.0 ? a * pow(q, b) : 0.0;
    const double phi = 1.0 / (1.0 + p);
    const double rest = p / (1.0 + p);
    const double eps = 1e-4;
    return -w * log(phi > eps ? phi : eps) - (1.0 - w) * log(rest > eps ? rest : eps);
}

__global__ __launch_bounds__(kObjThreads) void cross_entropy_kernel(
    const float *head, int64_t h, const float *tail, int64_t d, const int32_t *offsets,
    const int32_t *indices, const float *weights, double a, double b, double *partial) {
    __shared__ double s_sum[kObjThreads];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kObjThreads + threadIdx.x;
    double acc = 0.0;
    if (i < h) {
        for (int32_t e = offsets[i]; e < offsets[i + 1]; ++e) {
            const int64_t j = static_cast<uint32_t>(indices[e]);
            double q = 0.0;
            for (int64_t c = 0; c < d; ++c) {
                const double diff = static_cast<double>(head[i * d + c]) -
                                    static_cast<double>(tail[j * d + c]);
                q += diff * diff;
            }
            acc += cross_entropy_term(q, static_cast<double>(weights[e]), a, b);
        }
    }
    s_sum[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kObjThreads / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s_sum[0];
}

__global__ __launch_bounds__(kObjThreads) void sum_partials_kernel(const double *partial,
                                                                   int64_t count, double *out) {
    __shared__ double s_sum[kObjThreads];
    double acc = 0.0;
    // lane l adds partials l, l + 256, ... in order
    for (int64_t k = threadIdx.x; k < count; k += kObjThreads) acc += partial[k];
    s_sum[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kObjThreads / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s_sum[0];
}

// ------------------------------------------------------------- radix sort
// Stable LSD radix sort of (u32 key, i32 value) pairs, 8-bit digits, four
// passes. Per pass: per-tile digit histograms (a tile is 2048 keys), a scan
// that gives every (digit, tile) its output offset (digit-major, after all
// smaller digits), and a scatter. A tile's keys are taken wave-major (wave w
// owns keys 512w..512w+511 of the tile, in rows of 64); the scatter ranks
// each key among the equal digits before it -- lanes with the same digit are
// found with ballots over the nine digit bits (eight plus "valid") and
// counted into the wave's own digit counters row by row, then the four
// waves' counters are prefixed -- so equal keys keep their input order. The
// tile is then reordered in LDS by digit and written out in that order, so
// consecutive lanes write consecutive positions of a digit's run. No kernel
// waits on another workgroup. The digit totals of all passes are counted
// once, before pass 0, into a workspace array that the last pass clears.
constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;

__global__ __launch_bounds__(kSortThreads) void radix_zero_kernel(uint32_t *x, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kSortThreads + threadIdx.x;
    if (i < n) x[i] = 0u;
}

// Per-tile histogram of one key byte (with totals: also the key-byte totals
// of all four passes, accumulated into totals, which must be zero on entry).
__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const uint32_t *keys, int64_t n,
                                                                 int shift, int64_t tiles,
                                                                 uint32_t *hist,
                                                                 uint32_t *totals) {
    __shared__ uint32_t h[256];
    __shared__ uint32_t t4[4][256];
    const int tid = threadIdx.x;
    h[tid] = 0u;
    if (totals)
        for (int d = 0; d < 4; ++d) t4[d][tid] = 0u;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
    for (int i = 0; i < kSortItems; ++i) {
        const int64_t p = base + i * kSortThreads + tid;
        if (p < n) {
            const uint32_t k = keys[p];
            atomicAdd(&h[(k >> shift) & 255u], 1u);
            if (totals)
                for (int d = 0; d < 4; ++d) atomicAdd(&t4[d][(k >> (8 * d)) & 255u], 1u);
        }
    }
    __syncthreads();
    hist[static_cast<int64_t>(tid) * tiles + blockIdx.x] = h[tid];
    if (totals)
        for (int d = 0; d < 4; ++d)
            if (t4[d][tid]) atomicAdd(&totals[d * 256 + tid], t4[d][tid]);
}

__device__ inline uint32_t block_exclusive_scan_u32(uint32_t v, uint32_t *s_wave, uint32_t &total) {
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
    for (int w = 0; w < kSortThreads / 64; ++w) {
        if (w < wave) before += s_wave[w];
        total += s_wave[w];
    }
    __syncthreads();
    return before + incl - v;
}

// Workgroup d: offsets of digit d in every tile.
__global__ __launch_bounds__(kSortThreads) void radix_scan_kernel(uint32_t *hist, int64_t tiles,
                                                                 const uint32_t *totals) {
    __shared__ uint32_t s_wave[kSortThreads / 64];
    const int d = blockIdx.x, tid = threadIdx.x;
    uint32_t all;
    const uint32_t below = block_exclusive_scan_u32(totals[tid], s_wave, all);
    __shared__ uint32_t s_base;
    if (tid == d) s_base = below;
    __syncthreads();
    uint32_t carry = s_base;
    uint32_t *row = hist + static_cast<int64_t>(d) * tiles;
    for (int64_t c0 = 0; c0 < tiles; c0 += kSortThreads) {
        const int64_t t = c0 + tid;
        const uint32_t v = t < tiles ? row[t] : 0u;
        uint32_t chunk;
        const uint32_t ex = block_exclusive_scan_u32(v, s_wave, chunk);
        if (t < tiles) row[t] = carry + ex;
        carry += chunk;
    }
}

__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint32_t *keys_in, const int32_t *vals_in, uint32_t *keys_out, int32_t *vals_out,
    int64_t n, int shift, int64_t tiles, const uint32_t *offsets, int32_t *ranks_out,
    const float *table_src, float *table_out, uint32_t *clear_totals) {
    constexpr int WAVES = kSortThreads / 64;
    constexpr int PER_WAVE = kSortTile / WAVES;
    __shared__ uint32_t s_off[256];             // output position of the tile's run of each digit
    __shared__ uint32_t s_start[256];           // start of each digit in the tile sorted by digit
    __shared__ uint32_t s_wcnt[WAVES][256];     // per-wave digit counts, then their wave prefix
    __shared__ uint32_t s_wave[WAVES];
    __shared__ uint32_t s_key[kSortTile];
    __shared__ int32_t s_val[kSortTile];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
    const int cnt = static_cast<int>(min<int64_t>(kSortTile, n - base));
    if (clear_totals && blockIdx.x == 0)
        for (int t = tid; t < 1024; t += kSortThreads) clear_totals[t] = 0u;
    s_off[tid] = offsets[static_cast<int64_t>(tid) * tiles + blockIdx.x];
    for (int w = 0; w < WAVES; ++w) s_wcnt[w][tid] = 0u;
    uint32_t k[kSortItems];
    int32_t v[kSortItems];
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        const int local = wave * PER_WAVE + i * 64 + lane;
        k[i] = 0u;
        v[i] = 0;
        if (local < cnt) {
            k[i] = keys_in[base + local];
            v[i] = vals_in ? vals_in[base + local] : static_cast<int32_t>(base + local);
        }
    }
    __syncthreads();
    // rank of each key among the equal digits of its wave before it
    const uint64_t below_lane = (1ull << lane) - 1ull;
    uint32_t lrank[kSortItems];
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        const bool valid = wave * PER_WAVE + i * 64 + lane < cnt;
        const uint32_t digit = valid ? ((k[i] >> shift) & 255u) : 256u;
        uint64_t same = ~0ull;
#pragma unroll
        for (int bit = 0; bit < 9; ++bit) {
            const bool set = (digit >> bit) & 1u;
            const uint64_t ballot = __ballot(set);
            same &= set ? ballot : ~ballot;
        }
        const uint32_t r = static_cast<uint32_t>(__popcll(same & below_lane));
        const uint32_t before = valid ? s_wcnt[wave][digit] : 0u;
        lrank[i] = before + r;
        if (valid && r == 0u) s_wcnt[wave][digit] = before + static_cast<uint32_t>(__popcll(same));
    }
    __syncthreads();
    // digit tid: wave prefix of its counts, and its start in the digit-sorted tile
    uint32_t tot = 0u;
    for (int w = 0; w < WAVES; ++w) {
        const uint32_t c = s_wcnt[w][tid];
        s_wcnt[w][tid] = tot;
        tot += c;
    }
    uint32_t all;
    s_start[tid] = block_exclusive_scan_u32(tot, s_wave, all);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        if (wave * PER_WAVE + i * 64 + lane < cnt) {
            const uint32_t digit = (k[i] >> shift) & 255u;
            const uint32_t lp = s_start[digit] + s_wcnt[wave][digit] + lrank[i];
            s_key[lp] = k[i];
            s_val[lp] = v[i];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        const int idx = i * kSortThreads + tid;
        if (idx < cnt) {
            const uint32_t key = s_key[idx];
            const int32_t val = s_val[idx];
            const uint32_t digit = (key >> shift) & 255u;
            const uint32_t pos = s_off[digit] + (static_cast<uint32_t>(idx) - s_start[digit]);
            keys_out[pos] = key;
            vals_out[pos] = val;
            if (ranks_out) ranks_out[val] = static_cast<int32_t>(pos);
            if (table_out)
                reinterpret_cast<float2 *>(table_out)[pos] =
                    reinterpret_cast<const float2 *>(table_src)[val];
        }
    }
}

// ------------------------------------------------------------ projection
__device__ inline uint32_t radix_key(float v) {
    uint32_t bits = __float_as_uint(v);
    if (bits == 0x80000000u) bits = 0u;
    return (bits & 0x80000000u) ? ~bits : (bits ^ 0x80000000u);
}

__device__ inline float key_value(uint32_t key) {
    const uint32_t bits = (key & 0x80000000u) ? (key ^ 0x80000000u) : ~key;
    return __uint_as_float(bits);
}

__global__ void project_kernel(const float *x, int64_t rows, int64_t d, int64_t s0, int64_t s1,
                               const float *dir, float *out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    float sum = 0.0f;
    for (int64_t c = 0; c < d; ++c) sum += x[i * s0 + c * s1] * dir[c];
    out[i] = sum;
}

__global__ void project_keys_kernel(const float *x, int64_t rows, int64_t d, const float *dir,
                                    uint32_t *keys, float *proj, unsigned long long *status,
                                    unsigned long long tag) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    float sum = 0.0f;
    if (d == 2) {
        const float2 p = reinterpret_cast<const float2 *>(x)[i];
        sum += p.x * dir[0];
        sum += p.y * dir[1];
    } else {
        for (int64_t c = 0; c < d; ++c) sum += x[i * d + c] * dir[c];
    }
    if (proj) proj[i] = sum;
    if (keys) keys[i] = radix_key(sum);
    if (!isfinite(sum)) atomicMin(status, tag | static_cast<unsigned long long>(i));
}

// ------------------------------------------- single-launch radix passes
// The projection sort of rank-ordered fits: the same stable LSD radix sort
// (four 8-bit passes over (key, row) pairs), with one launch per pass and no
// histogram or scan launches. A pass kernel takes its tile id from a counter
// in the order its workgroups start, ranks the tile's keys among equal digits
// with ballots (as radix_scatter_kernel) and publishes the tile's digit
// counts with a ready bit. Its output offsets add the counts of the earlier
// tiles of its group of kOsGroup tiles and the totals of the earlier groups,
// which the last tile of each group publishes once it has its group's
// counts. Every tile and group total waited on belongs to a workgroup that
// started earlier and waits only on tiles before it, so each wait ends (a
// spin limit reports a stall instead of waiting forever).
// The key-byte totals of all four passes are counted by the projection
// kernel, which also clears the published counts and the tile counters; the
// last workgroup to finish the last pass clears the totals.
constexpr int kOsThreads = 512;
constexpr int kOsItems = 8;
constexpr int kOsTile = kOsThreads * kOsItems;
constexpr uint32_t kOsReady = 0x80000000u;
constexpr int kOsSpinLimit = 1 << 22;
constexpr int kOsGroup = 16;   // tiles per group (the group totals make the second level)

__device__ inline uint32_t os_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scans over the 256 digits of two vectors by threads 0..255 (every
// thread of the workgroup must call it); s_tmp holds 8 words.
__device__ inline void os_scan_digits(uint32_t a, uint32_t b, uint32_t &ea, uint32_t &eb,
                                      uint32_t *s_tmp) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t ia = a, ib = b;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t ua = __shfl_up(ia, off, 64), ub = __shfl_up(ib, off, 64);
        if (lane >= off) {
            ia += ua;
            ib += ub;
        }
    }
    if (tid < 256 && lane == 63) {
        s_tmp[wave] = ia;
        s_tmp[4 + wave] = ib;
    }
    __syncthreads();
    uint32_t ba = 0, bb = 0;
    for (int w = 0; w < 4; ++w)
        if (w < wave) {
            ba += s_tmp[w];
            bb += s_tmp[4 + w];
        }
    ea = ba + ia - a;
    eb = bb + ib - b;
    __syncthreads();
}

__global__ __launch_bounds__(kOsThreads) void radix_onesweep_kernel(
    const uint32_t *keys_in, const int32_t *vals_in, uint32_t *keys_out, int32_t *vals_out,
    int64_t n, int shift, int64_t tiles, const uint32_t *totals, uint32_t *counts,
    uint32_t *group_totals, uint32_t *tile_ctr, uint32_t *done_ctr, uint32_t *clear_totals,
    int32_t *err,
    int32_t *ranks_out, const float *table_src, float *table_out, const int32_t *csr_off,
    int32_t *seg_out, int32_t *deg_out, const float4 *rows) {
    constexpr int WAVES = kOsThreads / 64;
    constexpr int PER_WAVE = kOsTile / WAVES;
    __shared__ uint32_t s_off[256];            // output position of the tile's run of each digit
    __shared__ uint32_t s_start[256];          // start of each digit in the tile sorted by digit
    __shared__ uint32_t s_wcnt[WAVES][256];    // per-wave digit counts, then their wave prefix
    __shared__ uint32_t s_part[2][WAVES][256]; // per-wave sums: earlier tiles of the group, earlier groups
    __shared__ uint32_t s_tmp[8];
    __shared__ uint32_t s_key[kOsTile];
    __shared__ int32_t s_val[kOsTile];
    __shared__ uint32_t s_tile, s_last;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
    for (int i = tid; i < WAVES * 256; i += kOsThreads) (&s_wcnt[0][0])[i] = 0u;
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t base = tile * kOsTile;
    const int cnt = static_cast<int>(min<int64_t>(kOsTile, n - base));
    uint32_t k[kOsItems];
    int32_t v[kOsItems];
#pragma unroll
    for (int i = 0; i < kOsItems; ++i) {
        const int local = wave * PER_WAVE + i * 64 + lane;
        k[i] = 0u;
        v[i] = 0;
        if (local < cnt) {
            k[i] = keys_in[base + local];
            v[i] = vals_in ? vals_in[base + local] : static_cast<int32_t>(base + local);
        }
    }
    // rank of each key among the equal digits of its wave before it
    const uint64_t below_lane = (1ull << lane) - 1ull;
    uint32_t lrank[kOsItems];
#pragma unroll
    for (int i = 0; i < kOsItems; ++i) {
        const bool valid = wave * PER_WAVE + i * 64 + lane < cnt;
        const uint32_t digit = valid ? ((k[i] >> shift) & 255u) : 256u;
        uint64_t same = ~0ull;
#pragma unroll
        for (int bit = 0; bit < 9; ++bit) {
            const bool set = (digit >> bit) & 1u;
            const uint64_t ballot = __ballot(set);
            same &= set ? ballot : ~ballot;
        }
        const uint32_t r = static_cast<uint32_t>(__popcll(same & below_lane));
        const uint32_t before = valid ? s_wcnt[wave][digit] : 0u;
        lrank[i] = before + r;
        if (valid && r == 0u) s_wcnt[wave][digit] = before + static_cast<uint32_t>(__popcll(same));
    }
    __syncthreads();
    // digit tid: wave prefix of its counts and the tile's count, published
    uint32_t tot = 0u, gtot = 0u;
    if (tid < 256) {
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = s_wcnt[w][tid];
            s_wcnt[w][tid] = tot;
            tot += c;
        }
        __hip_atomic_store(counts + tile * 256 + tid, tot | kOsReady, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        gtot = totals[tid];
    }
    uint32_t start, gbase;
    os_scan_digits(tot, gtot, start, gbase, s_tmp);
    if (tid < 256) s_start[tid] = start;
    // Rows to add: the earlier tiles of the group (counts) and the earlier
    // group totals. Wave w adds rows w, w + WAVES, ... of a list; lane l
    // holds digits 4l..4l+3. The group's last tile publishes the group total
    // as soon as it has its group's counts, before it waits on earlier
    // groups, so group totals never wait on each other.
    const int64_t grp = tile / kOsGroup, pos_in_group = tile % kOsGroup;
    const bool publishes = pos_in_group == kOsGroup - 1;
    int spins = 0;
    // adds rows [0, n_own) of the group's counts and [0, n_grp) of the group
    // totals into s_part[0] / s_part[1] (per wave)
    auto add_rows = [&](int64_t n_own, int64_t n_grp) {
        uint32_t acc[2][4] = {{0u, 0u, 0u, 0u}, {0u, 0u, 0u, 0u}};
        const int64_t n_rows = n_own + n_grp;
        for (int64_t r0 = wave; r0 < n_rows; r0 += 4 * WAVES) {
            uint32_t c[4][4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t r = r0 + u * WAVES;
                const uint32_t *row = r < n_own ? counts + (grp * kOsGroup + r) * 256
                                                : group_totals + (r - n_own) * 256;
#pragma unroll
                for (int j = 0; j < 4; ++j) c[u][j] = r < n_rows ? os_load(row + 4 * lane + j) : kOsReady;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t r = r0 + u * WAVES;
                const int side = r < n_own ? 0 : 1;
                const uint32_t *row = r < n_own ? counts + (grp * kOsGroup + r) * 256
                                                : group_totals + (r - n_own) * 256;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    while (!(c[u][j] & kOsReady)) {
                        __builtin_amdgcn_s_sleep(1);
                        if (++spins > kOsSpinLimit) {
                            atomicOr(err, 2);
                            c[u][j] = kOsReady;
                            break;
                        }
                        c[u][j] = os_load(row + 4 * lane + j);
                    }
                    acc[side][j] += c[u][j] & ~kOsReady;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            s_part[0][wave][4 * lane + j] += acc[0][j];
            s_part[1][wave][4 * lane + j] += acc[1][j];
        }
    };
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        s_part[0][wave][4 * lane + j] = 0u;
        s_part[1][wave][4 * lane + j] = 0u;
    }
    if (publishes) {
        add_rows(pos_in_group, 0);
        __syncthreads();
        if (tid < 256) {
            uint32_t in_group = tot;
            for (int w = 0; w < WAVES; ++w) in_group += s_part[0][w][tid];
            __hip_atomic_store(group_totals + grp * 256 + tid, in_group | kOsReady,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        add_rows(0, grp);
    } else {
        add_rows(pos_in_group, grp);
    }
    __syncthreads();
    if (tid < 256) {
        uint32_t before = gbase;
        for (int w = 0; w < WAVES; ++w) before += s_part[0][w][tid] + s_part[1][w][tid];
        s_off[tid] = before;
    }
#pragma unroll
    for (int i = 0; i < kOsItems; ++i) {
        if (wave * PER_WAVE + i * 64 + lane < cnt) {
            const uint32_t digit = (k[i] >> shift) & 255u;
            const uint32_t lp = s_start[digit] + s_wcnt[wave][digit] + lrank[i];
            s_key[lp] = k[i];
            s_val[lp] = v[i];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kOsItems; ++i) {
        const int idx = i * kOsThreads + tid;
        if (idx < cnt) {
            const uint32_t key = s_key[idx];
            const int32_t val = s_val[idx];
            const uint32_t digit = (key >> shift) & 255u;
            const uint32_t pos = s_off[digit] + (static_cast<uint32_t>(idx) - s_start[digit]);
            keys_out[pos] = key;
            vals_out[pos] = val;
            if (ranks_out) ranks_out[val] = static_cast<int32_t>(pos);
            if (rows) {
                // one 16-byte gather per key
                const float4 r = rows[val];
                reinterpret_cast<float2 *>(table_out)[pos] = make_float2(r.x, r.y);
                seg_out[pos] = __float_as_int(r.z);
                deg_out[pos] = __float_as_int(r.w);
            } else if (table_out) {
                reinterpret_cast<float2 *>(table_out)[pos] =
                    reinterpret_cast<const float2 *>(table_src)[val];
            }
            if (seg_out && !rows) {
                const int32_t b = csr_off[val];
                seg_out[pos] = b;
                deg_out[pos] = csr_off[val + 1] - b;
            }
        }
    }
    if (clear_totals) {
        // every workgroup has read its totals before it counts itself done
        __syncthreads();
        if (tid == 0) s_last = atomicAdd(done_ctr, 1u) == static_cast<uint32_t>(tiles - 1);
        __syncthreads();
        if (s_last)
            for (int t = tid; t < 1024; t += kOsThreads) clear_totals[t] = 0u;
    }
}

// Keys of the d = 2 projection x . dir (the arithmetic of
// project_keys_kernel) and the key-byte totals of all four passes
// (totals must be zero on entry); also clears the published tile counts
// (`clear_words` words) and the five pass counters of the single-launch sort.
__global__ __launch_bounds__(kSortThreads) void project_keys_totals_kernel(
    const float *x, int64_t rows, const float *dir, uint32_t *keys, unsigned long long *status,
    unsigned long long tag, uint32_t *totals, uint32_t *clear, int64_t clear_words,
    uint32_t *ctr) {
    __shared__ uint32_t t4[4][256];
    const int tid = threadIdx.x;
    for (int d = 0; d < 4; ++d) t4[d][tid] = 0u;
    const int64_t per = (clear_words + gridDim.x - 1) / gridDim.x;
    const int64_t c0 = static_cast<int64_t>(blockIdx.x) * per;
    for (int64_t i = c0 + tid; i < min(c0 + per, clear_words); i += kSortThreads) clear[i] = 0u;
    if (blockIdx.x == 0 && tid < 5) ctr[tid] = 0u;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
    for (int i = 0; i < kSortItems; ++i) {
        const int64_t p = base + i * kSortThreads + tid;
        if (p < rows) {
            const float2 xy = reinterpret_cast<const float2 *>(x)[p];
            float sum = 0.0f;
            sum += xy.x * dir[0];
            sum += xy.y * dir[1];
            const uint32_t k = radix_key(sum);
            keys[p] = k;
            if (!isfinite(sum)) atomicMin(status, tag | static_cast<unsigned long long>(p));
            for (int d = 0; d < 4; ++d) atomicAdd(&t4[d][(k >> (8 * d)) & 255u], 1u);
        }
    }
    __syncthreads();
    for (int d = 0; d < 4; ++d)
        if (t4[d][tid]) atomicAdd(&totals[d * 256 + tid], t4[d][tid]);
}

__global__ void keys_from_values_kernel(const float *values, int64_t n, uint32_t *keys,
                                        unsigned long long *status) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = values[i];
    keys[i] = radix_key(v);
    if (!isfinite(v)) atomicMin(status, static_cast<unsigned long long>(i));
}

__global__ void iota_kernel(int32_t *out, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<int32_t>(i);
}

__global__ void values_from_keys_kernel(const uint32_t *keys, int64_t n, float *values) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) values[i] = key_value(keys[i]);
}

__global__ void inverse_ranks_kernel(const int32_t *order, int64_t n, int32_t *ranks) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < n) ranks[static_cast<uint32_t>(order[j])] = static_cast<int32_t>(j);
}

__global__ void lower_bound_kernel(const float *sorted, int64_t n, const float *query, int64_t h,
                                   int32_t *ranks) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= h) return;
    const float value = query[i];
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (sorted[mid] < value) lo = mid + 1;
        else hi = mid;
    }
    ranks[i] = static_cast<int32_t>(lo < n - 1 ? lo : n - 1);
}

__global__ void rank_and_table_kernel(const int32_t *order, int64_t n, const float *src,
                                      int32_t *ranks, float *table) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t id = static_cast<uint32_t>(order[r]);
    if (ranks) ranks[id] = static_cast<int32_t>(r);
    reinterpret_cast<float2 *>(table)[r] = reinterpret_cast<const float2 *>(src)[id];
}

__global__ void block_of_row_kernel(const int32_t *source_order, int64_t h, int64_t block,
                                    int32_t *block_of_row) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s < h) block_of_row[static_cast<uint32_t>(source_order[s])] = static_cast<int32_t>(s / block);
}

__global__ void momentum_apply_rows_kernel(int64_t h, float *head, const float *vel,
                                           const int32_t *block_of_row, int32_t block_id) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= h || block_of_row[i] != block_id) return;
    float2 x = reinterpret_cast<float2 *>(head)[i];
    const float2 v = reinterpret_cast<const float2 *>(vel)[i];
    x.x = x.x + v.x;
    x.y = x.y + v.y;
    reinterpret_cast<float2 *>(head)[i] = x;
}

__global__ void status_min_kernel(unsigned long long *status, unsigned long long value) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *status = value;
}

// ------------------------------------------------------------ adaptation
__device__ double resultant_factor(const float *emb, int64_t d, const int32_t *offsets,
                                   const int32_t *indices, int64_t src, int64_t j, double a,
                                   double b, double gamma, bool exclude) {
    if (exclude) {
        const int64_t begin = offsets[src], stop = offsets[src + 1];
        if (excluded(indices, begin, stop, src, j, true)) return 0.0;
    }
    double q = 0.0;
    for (int64_t c = 0; c < d; ++c) {
        const double z = static_cast<double>(emb[src * d + c]) - static_cast<double>(emb[j * d + c]);
        q += z * z;
    }
    const double coefficient = 2.0 * gamma * b / ((0.001 + q) * (a * pow(q, b) + 1.0));
    const double norm = coefficient * sqrt(q);
    if (norm > 0.0) return coefficient * gamma * tanh(norm / gamma) / norm;
    return 0.0;
}

__global__ void resultants_kernel(const float *emb, int64_t n, int64_t d, const int32_t *order,
                                  const int32_t *offsets, const int32_t *indices,
                                  const int32_t *sample_ranks, const int32_t *local_offsets,
                                  const int32_t *global_ranks, int64_t samples,
                                  uint32_t raw_range, float a, float b, float gamma, int exclude,
                                  double *out) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= 2 * samples) return;
    const bool local = t >= samples;
    const int64_t s = local ? t - samples : t;
    const int64_t src = static_cast<uint32_t>(order[static_cast<uint32_t>(sample_ranks[s])]);
    int64_t js[5];
    for (int k = 0; k < 5; ++k) {
        int64_t rank = static_cast<uint32_t>(global_ranks[s * 5 + k]);
        if (local) {
            const int64_t range = raw_range;
            const int64_t source_rank = static_cast<uint32_t>(sample_ranks[s]);
            const int64_t half = range / 2;
            int64_t first = source_rank > half ? source_rank - half : 0;
            if (first > n - range) first = n - range;
            rank = first + static_cast<uint32_t>(local_offsets[s * 5 + k]);
        }
        js[k] = static_cast<uint32_t>(order[rank]);
    }
    double f[5];
    for (int k = 0; k < 5; ++k)
        f[k] = resultant_factor(emb, d, offsets, indices, src, js[k], static_cast<double>(a),
                                static_cast<double>(b), static_cast<double>(gamma), exclude != 0);
    double squared = 0.0;
    for (int64_t c = 0; c < d; ++c) {
        const double x = static_cast<double>(emb[src * d + c]);
        double sum = f[0] * (x - static_cast<double>(emb[js[0] * d + c]));
        sum = sum + f[1] * (x - static_cast<double>(emb[js[1] * d + c]));
        sum = sum + f[2] * (x - static_cast<double>(emb[js[2] * d + c]));
        sum = sum + f[3] * (x - static_cast<double>(emb[js[3] * d + c]));
        sum = sum + f[4] * (x - static_cast<double>(emb[js[4] * d + c]));
        squared += sum * sum;
    }
    out[t] = sqrt(squared);
}

// negative_force_ratio + NegativeScaleAdaptation::observe_ratio
// (umap-core/src/modern_layout.rs:565-653) in one workgroup.
__global__ __launch_bounds__(256) void ratio_update_kernel(const double *res, int64_t samples,
                                                           double *state, float *scale_out,
                                                           int32_t *error_flag) {
    __shared__ double logs[kDeviceRatioMax];
    __shared__ int count_s;
    __shared__ int bad_s;
    const int tid = threadIdx.x;
    if (tid == 0) {
        // Sequential compaction keeps the valid pairs in sample order.
        int count = 0, bad = 0;
        for (int64_t s = 0; s < samples; ++s) {
            const double g = res[s], l = res[samples + s];
            if (!isfinite(g) || !isfinite(l) || g < 0.0 || l < 0.0) bad = 1;
            if (g > 0.0 && l > 0.0) logs[count++] = log(g) - log(l);
        }
        count_s = count;
        bad_s = bad;
    }
    __syncthreads();
    const int count = count_s;
    if (bad_s) {
        if (tid == 0) atomicOr(error_flag, 1);
        return;
    }
    // Bitonic sort of logs[0, pow2) padded with +inf.
    int p2 = 1;
    while (p2 < count) p2 <<= 1;
    for (int k = count + tid; k < p2; k += blockDim.x) logs[k] = INFINITY;
    __syncthreads();
    for (int size = 2; size <= p2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int k = tid; k < p2; k += blockDim.x) {
                const int partner = k ^ stride;
                if (partner > k) {
                    const bool up = (k & size) == 0;
                    const double x = logs[k], y = logs[partner];
                    if ((x > y) == up) {
                        logs[k] = y;
                        logs[partner] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (tid != 0) return;
    double ratio = 1.0;
    if (count > 0) {
        double qv[2];
        const double qs[2] = {0.1, 0.9};
        for (int k = 0; k < 2; ++k) {
            const double position = static_cast<double>(count - 1) * qs[k];
            const int lower = static_cast<int>(floor(position));
            const int upper = static_cast<int>(ceil(position));
            qv[k] = logs[lower] + (logs[upper] - logs[lower]) * (position - static_cast<double>(lower));
        }
        double sum = 0.0;
        for (int s = 0; s < count; ++s) {
            double v = logs[s];
            v = v < qv[0] ? qv[0] : (v > qv[1] ? qv[1] : v);
            sum += v;
        }
        ratio = exp(sum / static_cast<double>(count));
        if (!isfinite(ratio) || ratio <= 0.0) {
            atomicOr(error_flag, 1);
            return;
        }
    }
    // state: prior, current, anchor, has_anchor
    const double prior = state[0];
    double current = state[1];
    if (prior == 0.0) {
        state[2] = 0.0;
        state[3] = 1.0;
        *scale_out = 0.0f;
        return;
    }
    if (state[3] != 0.0) {
        double target = state[2] * ratio;
        const double lo = 0.5 * prior, hi = 2.0 * prior;
        target = target < lo ? lo : (target > hi ? hi : target);
        double change = log(target / current);
        const double clo = log(0.9), chi = log(1.1);
        change = change < clo ? clo : (change > chi ? chi : change);
        current *= exp(0.1 * change);
        state[1] = current;
    } else {
        const double anchor = current / ratio;
        if (!isfinite(anchor) || anchor <= 0.0) {
            atomicOr(error_flag, 1);
            return;
        }
        state[2] = anchor;
        state[3] = 1.0;
    }
    *scale_out = static_cast<float>(current);
}

}  // namespace

// ------------------------------------------------------------- launchers
void launch_gradient_rowwise(const RowwiseArgs &a, int64_t start, int64_t end, hipStream_t s) {
    if (end <= start) return;
    hipLaunchKernelGGL(gradient_rowwise_kernel, dim3(grid_for(end - start, 64)), dim3(64), 0, s,
                       a, start, end);
    HIP_CHECK(hipGetLastError());
}

void launch_adam(int64_t h, int64_t d, float *head, int64_t hs0, int64_t hs1, float *upd,
                 int64_t us0, int64_t us1, float *m, int64_t ms0, int64_t ms1, float *v,
                 int64_t vs0, int64_t vs1, float alpha, float c1, float c2, hipStream_t s) {
    if (h * d == 0) return;
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(h * d)), dim3(kThreads), 0, s, h, d, head, hs0,
                       hs1, upd, us0, us1, m, ms0, ms1, v, vs0, vs1, alpha, c1, c2);
    HIP_CHECK(hipGetLastError());
}

void launch_momentum_apply(int64_t d, float *head, int64_t hs0, int64_t hs1, const float *vel,
                           int64_t vs0, int64_t vs1, const int32_t *order, int64_t start,
                           int64_t end, hipStream_t s) {
    if (end <= start) return;
    hipLaunchKernelGGL(momentum_apply_kernel, dim3(grid_for(end - start)), dim3(kThreads), 0, s,
                       d, head, hs0, hs1, vel, vs0, vs1, order, start, end);
    HIP_CHECK(hipGetLastError());
}

void launch_momentum_decay(int64_t h, int64_t d, float *vel, int64_t vs0, int64_t vs1, float mu,
                           hipStream_t s) {
    if (h * d == 0) return;
    hipLaunchKernelGGL(momentum_decay_kernel, dim3(grid_for(h * d)), dim3(kThreads), 0, s, h, d,
                       vel, vs0, vs1, mu);
    HIP_CHECK(hipGetLastError());
}

void launch_edge_sweep(const SweepArgs &a, hipStream_t s) {
    if (a.h == 0) return;
    hipLaunchKernelGGL(edge_sweep_kernel, dim3(grid_for(a.h, kTileRows)), dim3(kThreads), 0, s,
                       a);
    HIP_CHECK(hipGetLastError());
}

int64_t unit_edges() { return kUnitEdges; }
int64_t unit_narrow_rows() { return kUnitNarrowRows; }

template <bool RANKED, int FMT, bool TRACE>
void launch_unit_sweep_rows(const UnitSweepArgs &a, dim3 grid, bool narrow, hipStream_t s) {
    if (narrow)
        hipLaunchKernelGGL((unit_sweep_kernel<RANKED, kUnitNarrowRows, FMT, TRACE>), grid, dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((unit_sweep_kernel<RANKED, kUnitEdges, FMT, TRACE>), grid, dim3(64), 0, s, a);
}

template <bool TRACE>
void launch_unit_sweep_shape(const UnitSweepArgs &a, dim3 grid, bool narrow, hipStream_t s) {
    if (a.pref && a.sched) launch_unit_sweep_rows<true, 2, TRACE>(a, grid, narrow, s);
    else if (a.pref && a.compact) launch_unit_sweep_rows<true, 1, TRACE>(a, grid, narrow, s);
    else if (a.pref) launch_unit_sweep_rows<true, 0, TRACE>(a, grid, narrow, s);
    else if (a.sched) launch_unit_sweep_rows<false, 2, TRACE>(a, grid, narrow, s);
    else launch_unit_sweep_rows<false, 0, TRACE>(a, grid, narrow, s);
}

void launch_unit_sweep(UnitSweepArgs a, hipStream_t s) {
    if (a.units == 0) return;
    const uint64_t modulus = a.s.range;
    a.t1 = (static_cast<uint64_t>(a.s.epoch) + 1u) % modulus;
    a.inv_mod = 1.0 / static_cast<double>(modulus);
    a.xcd_chunk = (a.units + 7) / 8;
    const dim3 grid(static_cast<unsigned>(8 * a.xcd_chunk));
    const bool narrow = a.row_cap == kUnitNarrowRows;
    UMAPHIP_REQUIRE(a.pref || !a.compact, UMAPHIP_BACKEND,
                    "unit sweep: compact records need the rank-ordered sweep");
    if (a.tr_count) {
        UMAPHIP_REQUIRE(a.tr_cap > 0 && a.tr_cand && a.tr_kept && a.tr_overflow && a.tr_paths,
                        UMAPHIP_INVALID_INPUT, "unit sweep trace buffers are null");
        launch_unit_sweep_shape<true>(a, grid, narrow, s);
    } else {
        launch_unit_sweep_shape<false>(a, grid, narrow, s);
    }
    HIP_CHECK(hipGetLastError());
}

int64_t rank_prefix_blocks(int64_t n) { return (n + kPrefBlock - 1) / kPrefBlock; }

void launch_rank_prefix(const int32_t *rank_deg, int64_t n, int32_t total_edges, int32_t *pref,
                        int32_t *block_sums, int32_t *unit_k, hipStream_t s) {
    if (n == 0 || total_edges == 0) return;
    const int64_t nb = rank_prefix_blocks(n);
    hipLaunchKernelGGL(rank_prefix_kernel, dim3(static_cast<unsigned>(nb)), dim3(kPrefThreads), 0,
                       s, rank_deg, n, pref, block_sums);
    hipLaunchKernelGGL(prefix_units_kernel, dim3(static_cast<unsigned>((n + kPrefThreads - 1) /
                                                                       kPrefThreads)),
                       dim3(kPrefThreads), 0, s, pref, n, block_sums, rank_deg, total_edges,
                       unit_k);
    HIP_CHECK(hipGetLastError());
}

void launch_unit_adam_ranked(int64_t h, const float *pos, float *pos_out, const float *grad,
                             uint8_t *split, const float *frag, float *m, float *v,
                             float alpha, float c1, float c2, hipStream_t s, float *grad_out,
                             const int32_t *csr_off, float *rows_out) {
    if (h == 0) return;
    UMAPHIP_REQUIRE(!rows_out || csr_off, UMAPHIP_INVALID_INPUT,
                    "packed rows need the CSR offsets");
    hipLaunchKernelGGL(unit_adam_ranked_kernel, dim3(grid_for(h)), dim3(kThreads), 0, s, h, pos,
                       pos_out, grad, split, frag, m, v, alpha, c1, c2, grad_out, csr_off,
                       reinterpret_cast<float4 *>(rows_out));
    HIP_CHECK(hipGetLastError());
}

void launch_pack_rows(const float *pos, const int32_t *csr_off, int64_t h, float *rows,
                      hipStream_t s) {
    if (h == 0) return;
    hipLaunchKernelGGL(pack_rows_kernel, dim3(grid_for(h)), dim3(kThreads), 0, s, pos, csr_off, h,
                       reinterpret_cast<float4 *>(rows));
    HIP_CHECK(hipGetLastError());
}

int64_t tile_rows(int shape) {
    switch (shape) {
        case 1: return 32;
        case 2: return 16;
        default: return kTileRows;
    }
}

void launch_segment_sweep(const SweepSegment *segs, int32_t n_segs, int64_t total_tiles,
                          hipStream_t s, int shape) {
    if (n_segs == 0 || total_tiles == 0) return;
    const dim3 grid(static_cast<unsigned>(total_tiles));
    switch (shape) {
        case 1:
            hipLaunchKernelGGL((segment_sweep_kernel<32, 128>), grid, dim3(128), 0, s, segs, n_segs);
            break;
        case 2:
            hipLaunchKernelGGL((segment_sweep_kernel<16, 64>), grid, dim3(64), 0, s, segs, n_segs);
            break;
        default:
            hipLaunchKernelGGL((segment_sweep_kernel<kTileRows, kThreads>), grid, dim3(kThreads), 0,
                               s, segs, n_segs);
            break;
    }
    HIP_CHECK(hipGetLastError());
}

__global__ void zero_i32_range_kernel(int32_t *p, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <int ROWS, int THREADS>
hipError_t launch_persistent_shape(const SweepSegment *segs, int32_t n_segs, int64_t total_tiles,
                                   int32_t epochs, int32_t *ctr, hipStream_t s) {
    auto kernel = persistent_segment_kernel<ROWS, THREADS>;
    int dev = 0, cus = 0, per_cu = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, THREADS, 0));
    UMAPHIP_REQUIRE(per_cu > 0, UMAPHIP_BACKEND, "persistent sweep: no resident workgroup fits");
    const int64_t blocks = std::min<int64_t>(total_tiles, static_cast<int64_t>(per_cu) * cus);
    const int64_t nctr = static_cast<int64_t>(epochs) * 8;
    hipLaunchKernelGGL(zero_i32_range_kernel, dim3(grid_for(nctr)), dim3(kThreads), 0, s, ctr, nctr);
    HIP_CHECK(hipGetLastError());
    void *args[] = {const_cast<SweepSegment **>(&segs), &n_segs, &total_tiles, &epochs, &ctr};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(kernel),
                                      dim3(static_cast<unsigned>(blocks)), dim3(THREADS), args, 0,
                                      s);
}

hipError_t launch_persistent_segments(const SweepSegment *segs, int32_t n_segs,
                                      int64_t total_tiles, int32_t epochs, int32_t *counters,
                                      hipStream_t s, int shape) {
    if (n_segs == 0 || total_tiles == 0 || epochs == 0) return hipSuccess;
    switch (shape) {
        case 1:
            return launch_persistent_shape<32, 128>(segs, n_segs, total_tiles, epochs, counters, s);
        case 2:
            return launch_persistent_shape<16, 64>(segs, n_segs, total_tiles, epochs, counters, s);
        default:
            return launch_persistent_shape<kTileRows, kThreads>(segs, n_segs, total_tiles, epochs,
                                                                counters, s);
    }
}

int64_t transform_tile_rows() { return kTransformRows; }

void launch_transform_layouts(const TransformSegment *segs, int32_t n_segs, int64_t total_tiles,
                              const TransformTables &tables, hipStream_t s) {
    if (n_segs == 0 || total_tiles == 0 || tables.epochs == 0) return;
    hipLaunchKernelGGL((transform_layout_kernel<kTransformRows, kTransformThreads>),
                       dim3(static_cast<unsigned>(total_tiles)), dim3(kTransformThreads), 0, s, segs,
                       n_segs, tables);
    HIP_CHECK(hipGetLastError());
}

int64_t cross_entropy_workspace(int64_t h) { return std::max<int64_t>((h + kObjThreads - 1) / kObjThreads, 1); }

void launch_cross_entropy(const float *head, int64_t h, const float *tail, int64_t d,
                          const int32_t *offsets, const int32_t *indices, const float *weights,
                          float a, float b, double *workspace, double *out, hipStream_t s) {
    const int64_t blocks = cross_entropy_workspace(h);
    if (h > 0) {
        hipLaunchKernelGGL(cross_entropy_kernel, dim3(static_cast<unsigned>(blocks)),
                           dim3(kObjThreads), 0, s, head, h, tail, d, offsets, indices, weights,
                           static_cast<double>(a), static_cast<double>(b), workspace);
        HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kObjThreads), 0, s, workspace,
                       h > 0 ? blocks : 0, out);
    HIP_CHECK(hipGetLastError());
}

void launch_unit_pack(const double *npos, const double *nneg, const double *periods,
                      const int32_t *indices, int64_t E, uint64_t modulus, bool compact,
                      double per_num, double *rec, uint16_t *fire, hipStream_t s) {
    if (E == 0) return;
    hipLaunchKernelGGL(unit_pack_kernel, dim3(grid_for(E)), dim3(kThreads), 0, s, npos, nneg,
                       periods, indices, E, modulus, compact, per_num, rec, fire);
    HIP_CHECK(hipGetLastError());
}

int64_t unit_schedule_stride(int64_t E) { return (E + 255) / 256 * 256; }

void launch_unit_schedule(double *rec, int64_t E, bool compact, double per_num, double rate,
                          bool negatives, int32_t t0, int32_t span, uint8_t *sched,
                          int64_t stride, int32_t *flag, hipStream_t s) {
    if (E == 0 || span <= 0) return;
    UMAPHIP_REQUIRE(stride % 4 == 0 && stride >= (E + 3) / 4 * 4, UMAPHIP_BACKEND,
                    "clock schedule: slab stride %lld does not hold %lld edges",
                    static_cast<long long>(stride), static_cast<long long>(E));
    hipLaunchKernelGGL(unit_schedule_kernel, dim3(grid_for((E + 3) / 4)), dim3(kThreads), 0, s,
                       rec, E, compact, per_num, rate, negatives, t0, span, sched, stride, flag);
    HIP_CHECK(hipGetLastError());
}

void launch_edge_schedule(double *npos, double *nneg, const double *periods,
                          const double *negative_periods, int64_t E, bool negatives, int32_t t0,
                          int32_t span, uint8_t *sched, int64_t stride, int32_t *flag,
                          hipStream_t s) {
    if (E == 0 || span <= 0) return;
    UMAPHIP_REQUIRE(stride % 4 == 0 && stride >= (E + 3) / 4 * 4, UMAPHIP_BACKEND,
                    "clock schedule: slab stride %lld does not hold %lld edges",
                    static_cast<long long>(stride), static_cast<long long>(E));
    hipLaunchKernelGGL(edge_schedule_kernel, dim3(grid_for((E + 3) / 4)), dim3(kThreads), 0, s,
                       npos, nneg, periods, negative_periods, E, negatives, t0, span, sched,
                       stride, flag);
    HIP_CHECK(hipGetLastError());
}

void launch_batch_table(const float *pos, const int32_t *orders, const int64_t *row_start,
                        const int64_t *order_start, const int32_t *graph_of_row,
                        int64_t total_rows, int32_t t, float *table, hipStream_t s) {
    if (total_rows == 0) return;
    hipLaunchKernelGGL(batch_table_kernel, dim3(grid_for(total_rows)), dim3(kThreads), 0, s, pos,
                       orders, row_start, order_start, graph_of_row, total_rows, t, table);
    HIP_CHECK(hipGetLastError());
}

void launch_unit_unpack(const double *rec, int64_t E, bool compact, double *npos, double *nneg,
                        hipStream_t s) {
    if (E == 0) return;
    hipLaunchKernelGGL(unit_unpack_kernel, dim3(grid_for(E)), dim3(kThreads), 0, s, rec, E, compact,
                       npos, nneg);
    HIP_CHECK(hipGetLastError());
}

bool periods_compactable(const double *periods, int64_t E, double per_num) {
    std::atomic<int> differs{0};
    parallel_for(E, [&](int64_t e) {
        const float samples = static_cast<float>(per_num / periods[e]);
        if (per_num / static_cast<double>(samples) != periods[e]) differs.store(1);
    }, 1 << 16);
    return !differs.load();
}

void launch_unit_adam(int64_t h, const float *pos, float *pos_out, const float *grad,
                      const int32_t *row_split, const int32_t *split_off, const float *frag,
                      float *m, float *v, float alpha, float c1, float c2, hipStream_t s,
                      float *grad_out) {
    if (h == 0) return;
    hipLaunchKernelGGL(unit_adam_kernel, dim3(grid_for(h)), dim3(kThreads), 0, s, h, pos, pos_out,
                       grad, row_split, split_off, frag, m, v, alpha, c1, c2, grad_out);
    HIP_CHECK(hipGetLastError());
}

void launch_debug_negatives(const RowwiseArgs &a, int32_t cap, int32_t *count_out,
                            int32_t *cand_out, uint8_t *kept_out, int32_t *overflow,
                            hipStream_t s) {
    if (a.h == 0) return;
    hipLaunchKernelGGL(debug_negatives_kernel, dim3(grid_for(a.h, 64)), dim3(64), 0, s, a, cap,
                       count_out, cand_out, kept_out, overflow);
    HIP_CHECK(hipGetLastError());
}

void launch_project(const float *x, int64_t rows, int64_t d, int64_t s0, int64_t s1,
                    const float *dir, float *out, hipStream_t s) {
    if (rows == 0) return;
    hipLaunchKernelGGL(project_kernel, dim3(grid_for(rows)), dim3(kThreads), 0, s, x, rows, d, s0,
                       s1, dir, out);
    HIP_CHECK(hipGetLastError());
}

void launch_project_keys(const float *x, int64_t rows, int64_t d, const float *dir,
                         uint32_t *keys, float *proj_out, unsigned long long *status,
                         unsigned long long tag, hipStream_t s) {
    if (rows == 0) return;
    hipLaunchKernelGGL(project_keys_kernel, dim3(grid_for(rows)), dim3(kThreads), 0, s, x, rows, d,
                       dir, keys, proj_out, status, tag);
    HIP_CHECK(hipGetLastError());
}

void launch_keys_from_values(const float *values, int64_t n, uint32_t *keys,
                             unsigned long long *status, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(keys_from_values_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, values, n,
                       keys, status);
    HIP_CHECK(hipGetLastError());
}

void launch_iota(int32_t *out, int64_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, out, n);
    HIP_CHECK(hipGetLastError());
}

int64_t sort_tiles(int64_t n) { return (n + kSortTile - 1) / kSortTile; }

size_t sort_pairs_workspace(int64_t n) {
    auto up = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
    return up(static_cast<size_t>(n) * 4) * 2 + up(static_cast<size_t>(256 * sort_tiles(n)) * 4) +
           up(1024 * 4);
}

namespace {
struct SortWorkspace {
    uint32_t *tmp_k;
    int32_t *tmp_v;
    uint32_t *hist, *totals;
    int64_t tiles;
};
SortWorkspace sort_workspace(void *ws, size_t ws_bytes, int64_t n) {
    UMAPHIP_REQUIRE(n <= INT32_MAX, UMAPHIP_OVERFLOW, "sort size must fit i32");
    UMAPHIP_REQUIRE(ws_bytes >= sort_pairs_workspace(n), UMAPHIP_CAPACITY, "sort workspace too small");
    auto up = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
    uint8_t *w = static_cast<uint8_t *>(ws);
    SortWorkspace W;
    W.tiles = sort_tiles(n);
    W.tmp_k = reinterpret_cast<uint32_t *>(w);
    W.tmp_v = reinterpret_cast<int32_t *>(w + up(n * 4));
    W.hist = reinterpret_cast<uint32_t *>(w + 2 * up(n * 4));
    W.totals = reinterpret_cast<uint32_t *>(w + 2 * up(n * 4) + up(256 * W.tiles * 4));
    return W;
}

// The four passes: in -> tmp -> out -> tmp -> out; the last pass clears the
// totals.
void sort_passes(const SortWorkspace &W, const uint32_t *keys_in, uint32_t *keys_out,
                 const int32_t *vals_in, int32_t *vals_out, int64_t n, int32_t *ranks,
                 const float *table_src, float *table_out, hipStream_t s) {
    const unsigned tiles = static_cast<unsigned>(W.tiles);
    const uint32_t *src_k = keys_in;
    const int32_t *src_v = vals_in;
    for (int pass = 0; pass < 4; ++pass) {
        uint32_t *dst_k = (pass & 1) ? keys_out : W.tmp_k;
        int32_t *dst_v = (pass & 1) ? vals_out : W.tmp_v;
        const int shift = 8 * pass;
        hipLaunchKernelGGL(radix_hist_kernel, dim3(tiles), dim3(kSortThreads), 0, s, src_k, n,
                           shift, W.tiles, W.hist, pass == 0 ? W.totals : nullptr);
        hipLaunchKernelGGL(radix_scan_kernel, dim3(256), dim3(kSortThreads), 0, s, W.hist, W.tiles,
                           W.totals + pass * 256);
        const bool last = pass == 3;
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(tiles), dim3(kSortThreads), 0, s, src_k, src_v,
                           dst_k, dst_v, n, shift, W.tiles, W.hist, last ? ranks : nullptr,
                           last ? table_src : nullptr, last ? table_out : nullptr,
                           last ? W.totals : nullptr);
        src_k = dst_k;
        src_v = dst_v;
    }
    HIP_CHECK(hipGetLastError());
}
}  // namespace

// With ranks (and table_src / table_out for d = 2), the last pass also
// writes ranks[value] = position and table_out[position] = table_src[value].
void launch_sort_pairs_ranked(void *ws, size_t ws_bytes, const uint32_t *keys_in,
                              uint32_t *keys_out, const int32_t *vals_in, int32_t *vals_out,
                              int64_t n, int32_t *ranks, const float *table_src, float *table_out,
                              hipStream_t s) {
    if (n == 0) return;
    const SortWorkspace W = sort_workspace(ws, ws_bytes, n);
    hipLaunchKernelGGL(radix_zero_kernel, dim3(4), dim3(kSortThreads), 0, s, W.totals, int64_t(1024));
    sort_passes(W, keys_in, keys_out, vals_in, vals_out, n, ranks, table_src, table_out, s);
}

int64_t onesweep_tiles(int64_t n) { return (n + kOsTile - 1) / kOsTile; }
int64_t onesweep_groups(int64_t n) { return (onesweep_tiles(n) + kOsGroup - 1) / kOsGroup; }

// per pass: tile counts [tiles][256], then group totals [groups][256]
int64_t onesweep_count_words(int64_t n) {
    return 4 * 256 * (onesweep_tiles(n) + onesweep_groups(n));
}

size_t project_sort_workspace(int64_t n) {
    auto up = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
    return up(static_cast<size_t>(n) * 4) * 2 +
           up(static_cast<size_t>(onesweep_count_words(n)) * 4) + up(1024 * 4) + up(8 * 4);
}

namespace {
struct OnesweepWorkspace {
    uint32_t *tmp_k;
    int32_t *tmp_v;
    uint32_t *counts, *group_totals, *totals, *ctr;
    int64_t tiles, groups;
};
OnesweepWorkspace onesweep_workspace(void *ws, size_t ws_bytes, int64_t n) {
    // published counts carry a ready bit in bit 31
    UMAPHIP_REQUIRE(n < (int64_t(1) << 31), UMAPHIP_OVERFLOW, "sort size must be below 2^31");
    UMAPHIP_REQUIRE(ws_bytes >= project_sort_workspace(n), UMAPHIP_CAPACITY,
                    "projection sort workspace too small");
    auto up = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
    uint8_t *w = static_cast<uint8_t *>(ws);
    OnesweepWorkspace W;
    W.tiles = onesweep_tiles(n);
    W.groups = onesweep_groups(n);
    const size_t kb = up(static_cast<size_t>(n) * 4);
    const size_t cb = up(static_cast<size_t>(onesweep_count_words(n)) * 4);
    W.tmp_k = reinterpret_cast<uint32_t *>(w);
    W.tmp_v = reinterpret_cast<int32_t *>(w + kb);
    W.counts = reinterpret_cast<uint32_t *>(w + 2 * kb);
    W.group_totals = W.counts + 4 * 256 * W.tiles;
    W.totals = reinterpret_cast<uint32_t *>(w + 2 * kb + cb);
    W.ctr = reinterpret_cast<uint32_t *>(w + 2 * kb + cb + up(1024 * 4));
    return W;
}
}  // namespace

void clear_project_sort_workspace(void *ws, size_t ws_bytes, int64_t n, hipStream_t s) {
    if (n == 0) return;
    const OnesweepWorkspace W = onesweep_workspace(ws, ws_bytes, n);
    hipLaunchKernelGGL(radix_zero_kernel, dim3(4), dim3(kSortThreads), 0, s, W.totals, int64_t(1024));
    HIP_CHECK(hipGetLastError());
}

void launch_project_sort_ranked(void *ws, size_t ws_bytes, const float *x, int64_t n,
                                const float *dir, unsigned long long *status,
                                unsigned long long tag, uint32_t *keys, uint32_t *keys_out,
                                int32_t *vals_out, int32_t *ranks, float *table_out, hipStream_t s,
                                const int32_t *csr_off, int32_t *seg_out, int32_t *deg_out,
                                int32_t *err, const float *rows) {
    if (n == 0) return;
    UMAPHIP_REQUIRE(err != nullptr, UMAPHIP_INVALID_INPUT, "projection sort needs an error word");
    UMAPHIP_REQUIRE(!rows || (table_out && seg_out && deg_out), UMAPHIP_INVALID_INPUT,
                    "packed rows feed the table and the rank-ordered CSR starts and degrees");
    const OnesweepWorkspace W = onesweep_workspace(ws, ws_bytes, n);
    const int64_t count_words = onesweep_count_words(n);
    hipLaunchKernelGGL(project_keys_totals_kernel, dim3(static_cast<unsigned>(sort_tiles(n))),
                       dim3(kSortThreads), 0, s, x, n, dir, keys, status, tag, W.totals, W.counts,
                       count_words, W.ctr);
    const uint32_t *src_k = keys;
    const int32_t *src_v = nullptr;
    for (int pass = 0; pass < 4; ++pass) {
        uint32_t *dst_k = (pass & 1) ? keys_out : W.tmp_k;
        int32_t *dst_v = (pass & 1) ? vals_out : W.tmp_v;
        const bool last = pass == 3;
        hipLaunchKernelGGL(radix_onesweep_kernel, dim3(static_cast<unsigned>(W.tiles)),
                           dim3(kOsThreads), 0, s, src_k, src_v, dst_k, dst_v, n, 8 * pass,
                           W.tiles, W.totals + 256 * pass, W.counts + 256 * W.tiles * pass,
                           W.group_totals + 256 * W.groups * pass, W.ctr + pass, W.ctr + 4,
                           last ? W.totals : nullptr, err,
                           last ? ranks : nullptr, last ? x : nullptr, last ? table_out : nullptr,
                           last ? csr_off : nullptr, last ? seg_out : nullptr,
                           last ? deg_out : nullptr,
                           last ? reinterpret_cast<const float4 *>(rows) : nullptr);
        src_k = dst_k;
        src_v = dst_v;
    }
    HIP_CHECK(hipGetLastError());
}

void launch_sort_pairs(void *ws, size_t ws_bytes, const uint32_t *keys_in, uint32_t *keys_out,
                       const int32_t *vals_in, int32_t *vals_out, int64_t n, hipStream_t s) {
    launch_sort_pairs_ranked(ws, ws_bytes, keys_in, keys_out, vals_in, vals_out, n, nullptr,
                             nullptr, nullptr, s);
}

void launch_values_from_keys(const uint32_t *keys, int64_t n, float *values, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(values_from_keys_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, keys, n,
                       values);
    HIP_CHECK(hipGetLastError());
}

void launch_inverse_ranks(const int32_t *order, int64_t n, int32_t *ranks, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(inverse_ranks_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, order, n,
                       ranks);
    HIP_CHECK(hipGetLastError());
}

void launch_lower_bound(const float *sorted, int64_t n, const float *query, int64_t h,
                        int32_t *ranks, hipStream_t s) {
    if (h == 0) return;
    hipLaunchKernelGGL(lower_bound_kernel, dim3(grid_for(h)), dim3(kThreads), 0, s, sorted, n,
                       query, h, ranks);
    HIP_CHECK(hipGetLastError());
}

void launch_rank_and_table(const int32_t *order, int64_t n, const float *src, int32_t *ranks,
                           float *table, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(rank_and_table_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, order, n,
                       src, ranks, table);
    HIP_CHECK(hipGetLastError());
}

void launch_gather_table(const int32_t *order, int64_t n, const float *src, float *table,
                         hipStream_t s) {
    launch_rank_and_table(order, n, src, nullptr, table, s);
}

void launch_block_of_row(const int32_t *source_order, int64_t h, int64_t block,
                         int32_t *block_of_row, hipStream_t s) {
    if (h == 0) return;
    hipLaunchKernelGGL(block_of_row_kernel, dim3(grid_for(h)), dim3(kThreads), 0, s, source_order,
                       h, block, block_of_row);
    HIP_CHECK(hipGetLastError());
}

void launch_momentum_apply_rows(int64_t h, float *head, const float *vel,
                                const int32_t *block_of_row, int32_t block_id, hipStream_t s) {
    if (h == 0) return;
    hipLaunchKernelGGL(momentum_apply_rows_kernel, dim3(grid_for(h)), dim3(kThreads), 0, s, h,
                       head, vel, block_of_row, block_id);
    HIP_CHECK(hipGetLastError());
}

void launch_status_min_u64(unsigned long long *status, unsigned long long value, hipStream_t s) {
    hipLaunchKernelGGL(status_min_kernel, dim3(1), dim3(64), 0, s, status, value);
    HIP_CHECK(hipGetLastError());
}

void launch_resultants(const float *emb, int64_t n, int64_t d, const int32_t *order,
                       const int32_t *offsets, const int32_t *indices, const int32_t *sample_ranks,
                       const int32_t *local_offsets, const int32_t *global_ranks, int64_t samples,
                       uint32_t raw_range, float a, float b, float gamma, int exclude,
                       double *out, hipStream_t s) {
    if (samples == 0) return;
    hipLaunchKernelGGL(resultants_kernel, dim3(grid_for(2 * samples, 64)), dim3(64), 0, s, emb, n,
                       d, order, offsets, indices, sample_ranks, local_offsets, global_ranks,
                       samples, raw_range, a, b, gamma, exclude, out);
    HIP_CHECK(hipGetLastError());
}

void launch_ratio_update(const double *resultants, int64_t samples, double *state,
                         float *scale_out, int32_t *error_flag, hipStream_t s) {
    hipLaunchKernelGGL(ratio_update_kernel, dim3(1), dim3(256), 0, s, resultants, samples, state,
                       scale_out, error_flag);
    HIP_CHECK(hipGetLastError());
}

}  // namespace umaphip
