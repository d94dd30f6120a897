// Exact k-nearest neighbours for the transform path and the fit graph:
// a BF16 MFMA distance prefilter (FP32 accumulation) that keeps the best 64
// candidates per query, then an exact FP32 rerank that reproduces cuscan's
// exact_kernel arithmetic (crates/cuscan-common/src/kernels/neighbors.rs:
// 17-76): sequential f32 sum of squared differences over the features,
// ties broken by the lower reference index, square root at the end,
// sentinel id n_reference and distance sqrt(f32::MAX) for unfilled slots.
//
// Prefilter layout (gfx950, wave64): a workgroup of 4 waves owns 128 queries
// (32 per wave, held for the whole scan as MFMA A-fragments in VGPRs) and
// streams one slice of the references through a double-buffered LDS stage
// in chunks of 16 rows: the next chunk is loaded straight from global
// memory into LDS (global_load_lds, one instruction per MFMA K step of the
// chunk) while the current one feeds the MFMAs.
// Approximate distances ||r||^2 - 2 q.r are compared against each query's
// running 64th-best threshold (kept in LDS); survivors are appended to
// a per-(query, slice) buffer that is compacted to its sorted best 64 when
// it could overflow. A wave owns the candidate state of its own 32 queries,
// so the threshold epilogue needs no workgroup barrier. Slices are merged
// per query, then reranked.
//
// Sample bound (batches with many slices): a slice that starts from an
// infinite threshold keeps its first 128 references and sorts its buffer
// about every 64 survivors until its threshold tightens, and a wave does
// that one query at a time, so small batches spent most of the scan there.
// Before such a scan, the approximate distances of every query to a sample
// of the references (8,192, or 32,768 for batches up to 4,096 rows, in
// blocks of 16 consecutive rows spread evenly over the references) are
// computed by the scan kernel itself, and each query's 64th best of them
// seeds its global threshold: 64 real references lie at or below it, so it
// bounds the query's final 64th best from above and no candidate of the
// exact top 64 is dropped.
//
// Centring and the certificate: the BF16 copies hold x - s for a per-feature
// shift s near the references' mean (distances do not change), so the
// rounding error is relative to the spread of the data, not to its offset.
// After the rerank, each query's result is certified: a bound on the BF16
// rounding (measured per row) and on the f32 accumulation of the
// approximate distances turns the 64th approximate distance into a lower
// bound on the exact distance of every reference left out of the 64; when
// that bound exceeds the reranked k-th distance, no left-out reference can
// enter the top k. Queries the bound does not certify are rescanned exactly
// in FP32 over all references, in the reference's summation order.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>
#include <type_traits>

#include "common.hpp"
#include "knn.hpp"
#include "wave.hpp"

namespace umaphip {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int kTQ = 128;      // queries per workgroup
constexpr int kTR = 16;       // reference rows per LDS chunk (one MFMA column block)
constexpr int kCap = 128;     // candidate buffer per (query, slice)
constexpr int kTop = 64;      // candidates kept per query
constexpr int kMaxSlices = 512;   // slices per query batch (KnnIndex::slices_for)
constexpr int64_t kMinSliceRefs = 2048;   // fewest references per slice
constexpr int kSample = 8192;      // references in the sample bound
constexpr int kSampleLarge = 32768;   // and for batches of at most kLargeSampleMaxQ rows
constexpr int64_t kLargeSampleMaxQ = 4096;
constexpr int kSelectKeys = 32;       // sample keys per thread of the select
constexpr int kSampleRefs = 16;    // fewest sample references per workgroup of the sample distances
constexpr int kSampleMinSlices = 8;   // batches with at least this many slices use the bound
constexpr int64_t kSampleMinRefs = 8 * kSample;
constexpr int64_t kWideMergeMaxQ = 8192;   // batches up to this many rows merge with 4 waves
constexpr int64_t kDeepMergeMaxQ = 512;    // and up to this many with 16
constexpr int kMergeAhead = 4;             // slice lists in flight per merging wave
constexpr int kRerankSlots = 16;           // features per lane per row of the wide rerank
constexpr int64_t kRerankWideMaxQ = 512;   // batches up to this many rows rerank wide
constexpr int kMomentRows = 2048;          // rows per workgroup of the column moments
constexpr int kExactQB = 4;                // queries per workgroup of the exact rescan
constexpr int64_t kExactRound = 1024;      // uncertified queries per rescan round
constexpr int kScanRing = 4;               // B fragments in flight in the large-grid scan

__device__ inline uint16_t f32_to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// BF16 rows of x - shift, padded to dp features; norms[r] = the f32 sum of
// the squared BF16 values, err[r] (may be null) = an upper bound on the
// Euclidean norm of (BF16 row) - (x_r - shift), the row's total rounding.
__global__ void to_bf16_kernel(const float *x, int64_t rows, int64_t d, int64_t dp,
                               const float *shift, uint16_t *out, float *norms, float *err) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.y + threadIdx.y;
    if (r >= rows) return;
    float acc = 0.0f;
    double esq = 0.0;
    for (int64_t c = threadIdx.x; c < dp; c += blockDim.x) {
        uint16_t h = 0;
        if (c < d) {
            const float xv = x[r * d + c], sv = shift[c];
            h = f32_to_bf16_rne(xv - sv);
            const double e = static_cast<double>(bf16_to_f32(h)) -
                             (static_cast<double>(xv) - static_cast<double>(sv));
            esq += e * e;
        }
        out[r * dp + c] = h;
        const float v = bf16_to_f32(h);
        acc += v * v;
    }
    for (int off = 32; off > 0; off >>= 1) {
        acc += __shfl_xor(acc, off, 64);
        esq += __shfl_xor(esq, off, 64);
    }
    if (threadIdx.x == 0) {
        if (norms) norms[r] = acc;
        if (err) err[r] = static_cast<float>(sqrt(esq) * (1.0 + 0x1.0p-20));
    }
}

// Per-feature f64 sums of x and x^2 over kMomentRows rows per workgroup:
// partial[(b * 2 + 0) * d + c] and partial[(b * 2 + 1) * d + c].
__global__ __launch_bounds__(256) void column_moments_kernel(const float *x, int64_t n, int64_t d,
                                                             double *partial) {
    const int64_t b = blockIdx.x;
    const int64_t r0 = b * kMomentRows, r1 = min<int64_t>(n, r0 + kMomentRows);
    for (int64_t c = threadIdx.x; c < d; c += blockDim.x) {
        double s1 = 0.0, s2 = 0.0;
        for (int64_t r = r0; r < r1; ++r) {
            const double v = x[r * d + c];
            s1 += v;
            s2 += v * v;
        }
        partial[(b * 2) * d + c] = s1;
        partial[(b * 2 + 1) * d + c] = s2;
    }
}

// The per-feature shift: the mean, rounded to a multiple of the largest
// power of two not above the feature's standard deviation, so it is exact
// in f32 and data on an integer grid keep their grid (0..255 data stay
// integers within +-255, exact in BF16). Partial sums are added in a fixed
// order, so the shift does not depend on scheduling.
__global__ void shift_kernel(const double *partial, int64_t blocks, int64_t n, int64_t d,
                             float *shift) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= d) return;
    double s1 = 0.0, s2 = 0.0;
    for (int64_t b = 0; b < blocks; ++b) {
        s1 += partial[(b * 2) * d + c];
        s2 += partial[(b * 2 + 1) * d + c];
    }
    const double mean = s1 / static_cast<double>(n);
    const double var = fmax(s2 / static_cast<double>(n) - mean * mean, 0.0);
    const double sd = sqrt(var);
    double v = static_cast<double>(static_cast<float>(mean));
    if (sd > 0.0 && isfinite(sd) && isfinite(mean)) {
        const double step = exp2(floor(log2(sd)));
        v = nearbyint(mean / step) * step;
    }
    shift[c] = isfinite(v) ? static_cast<float>(v) : 0.0f;
}

// eps_key = the order-preserving key of max over rows of err / ||BF16 row||:
// every reference row's rounding is at most eps times its BF16 norm (a row
// with rounding but zero BF16 norm makes eps infinite).
__global__ void rounding_ratio_kernel(const float *norms, const float *err, int64_t n, double g,
                                      uint32_t *eps_key) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    float ratio = 0.0f;
    if (r < n && err[r] > 0.0f) {
        const double nr = static_cast<double>(norms[r]) * (1.0 - g);
        ratio = nr > 0.0 ? static_cast<float>(static_cast<double>(err[r]) / sqrt(nr) *
                                              (1.0 + 0x1.0p-20))
                         : INFINITY;
    }
    for (int off = 32; off > 0; off >>= 1) ratio = fmaxf(ratio, __shfl_xor(ratio, off, 64));
    if ((threadIdx.x & 63) == 0 && ratio > 0.0f) atomicMax(eps_key, __float_as_uint(ratio));
}

// Sorts the n <= kCap = 128 entries of one (query, slice) buffer by
// (distance, id) with a bitonic network (two entries per lane, entry e in
// lane e % 64, register e / 64), keeps the best kTop = 64 sorted in place
// and returns the new threshold (the kTop-th distance, or +inf when fewer).
// Ids are unique, so the order is total.
__device__ inline void load_buffer(const float *cd, const int32_t *ci, int n, int lane,
                                   float (&d)[2], int32_t (&id)[2]) {
    d[0] = lane < n ? cd[lane] : INFINITY;
    id[0] = lane < n ? ci[lane] : INT32_MAX;
    d[1] = lane + 64 < n ? cd[lane + 64] : INFINITY;
    id[1] = lane + 64 < n ? ci[lane + 64] : INT32_MAX;
}

__device__ inline float sort_buffer(float (&d)[2], int32_t (&id)[2], int n, float *cd, int32_t *ci,
                                    int lane) {
    static_assert(kCap == 128 && kTop == 64, "two entries per lane");
#pragma unroll
    for (int k = 2; k <= 128; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j == 64) {
                // entries lane and lane + 64 (k = 128: ascending)
                const bool gt = d[0] > d[1] || (d[0] == d[1] && id[0] > id[1]);
                if (gt) {
                    const float td = d[0];
                    const int32_t ti = id[0];
                    d[0] = d[1];
                    id[0] = id[1];
                    d[1] = td;
                    id[1] = ti;
                }
                continue;
            }
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const float dp = __shfl_xor(d[r], j, 64);
                const int32_t ip = __shfl_xor(id[r], j, 64);
                const int e = lane + 64 * r;
                const bool lower = (lane & j) == 0;     // e is the lower entry of its pair
                const bool asc = (e & k) == 0;          // the pair sorts ascending
                const bool mine_gt = d[r] > dp || (d[r] == dp && id[r] > ip);
                // lower keeps the min when ascending, the max when descending
                const bool take = lower == asc ? mine_gt : !mine_gt;
                if (take) {
                    d[r] = dp;
                    id[r] = ip;
                }
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    cd[lane] = d[0];
    ci[lane] = id[0];
    const float kth = __shfl(d[0], kTop - 1, 64);
    return n >= kTop ? kth : INFINITY;
}

// sort_buffer for n <= 64 entries, one per lane (lanes n.. hold +inf): a
// bitonic network over the smallest power of two >= n lanes only, since the
// final buffers of small batches hold a few entries each. Writes the same
// 64 entries as sort_buffer.
__device__ inline void sort_short(float d, int32_t id, int n, float *cd, int32_t *ci, int lane) {
    int size = 2;
    while (size < n) size <<= 1;
    for (int k = 2; k <= size; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            const float dp = __shfl_xor(d, j, 64);
            const int32_t ip = __shfl_xor(id, j, 64);
            const bool lower = (lane & j) == 0;     // lane is the lower entry of its pair
            const bool asc = (lane & k) == 0;       // the pair sorts ascending
            const bool mine_gt = d > dp || (d == dp && id > ip);
            if (lower == asc ? mine_gt : !mine_gt) {
                d = dp;
                id = ip;
            }
        }
    cd[lane] = d;
    ci[lane] = id;
}

__device__ inline float compact_buffer(float *cd, int32_t *ci, int n, int lane) {
    float d[2];
    int32_t id[2];
    load_buffer(cd, ci, n, lane, d, id);
    return sort_buffer(d, id, n, cd, ci, lane);
}

// End of a scan cell: the final sort of the buffers of the block's kTQ
// queries and their counts (at most kTop). Query ql goes to wave ql % 4, so
// a batch of fewer than kTQ rows spreads over all four waves, and each wave
// loads its next query's buffer while it sorts the current one. Call after
// every wave's appends (global and s_cnt) are visible to the block.
__device__ inline void finish_cell(const int *s_cnt, int64_t q_block, int64_t nq, int slices,
                                   int slice, float *cand_d, int32_t *cand_i, int32_t *cand_n,
                                   int wave, int lane, int bq = kTQ) {
    int ql = wave;
    int64_t q = q_block + ql;
    int n = q < nq ? s_cnt[ql] : 0;
    float nd[2], d[2];
    int32_t nid[2], id[2];
    if (q < nq) load_buffer(cand_d + (q * slices + slice) * kCap, cand_i + (q * slices + slice) * kCap,
                            n, lane, nd, nid);
    while (q < nq) {
        const int cur_n = n;
        const int64_t cur = q * slices + slice;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            d[r] = nd[r];
            id[r] = nid[r];
        }
        ql += 4;
        q += 4;
        if (ql >= bq) q = nq;   // past the block
        if (q < nq) {
            n = s_cnt[ql];
            load_buffer(cand_d + (q * slices + slice) * kCap, cand_i + (q * slices + slice) * kCap,
                        n, lane, nd, nid);
        }
        if (cur_n > kTop)
            sort_buffer(d, id, cur_n, cand_d + cur * kCap, cand_i + cur * kCap, lane);
        else if (cur_n > 0)
            sort_short(d[0], id[0], cur_n, cand_d + cur * kCap, cand_i + cur * kCap, lane);
        if (lane == 0) cand_n[cur] = cur_n < kTop ? cur_n : kTop;
    }
}

// Order-preserving u32 key of a float (the global threshold is kept as a
// key so that atomicMin orders it); +inf maps to 0xFF800000.
__device__ inline uint32_t float_key(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float key_float(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__global__ void fill_u32_kernel(uint32_t *p, int64_t n, uint32_t v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Register staging of one chunk (the LDS image word j * 64 + lane of the
// padded rows, for this wave's j = wave, wave + 4, ...; past the slice a
// harmless in-bounds word), and its store into an LDS buffer.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// s_waitcnt vmcnt(N) (expcnt and lgkmcnt not waited on).
template <int N>
__device__ inline void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
    __builtin_amdgcn_s_waitcnt((N & 15) | (0x7 << 4) | (0xF << 8) | ((N >> 4) << 14));
}

template <int LW, int INSTR, int SLOTS_ROW, int DP>
__device__ __forceinline__ void fetch_chunk(u32x4 (&dst)[LW], float &rn, const uint16_t *rb,
                                            const float *rnorm, int64_t chunk, int64_t r_end,
                                            int wave, int lane) {
#pragma unroll
    for (int t = 0; t < LW; ++t) {
        const int j = wave + 4 * t;
        const int sl = j * 64 + lane;
        const int row = sl / SLOTS_ROW, col = sl % SLOTS_ROW;
        const int64_t ref = chunk + row;
        const uint16_t *src = rb;
        if (j < INSTR && row < kTR && col < DP / 8 && ref < r_end) src = rb + ref * DP + col * 8;
        dst[t] = *reinterpret_cast<const u32x4 *>(src);
    }
    rn = rnorm[min<int64_t>(chunk + (lane & 15), r_end - 1)];
}

template <int LW, int INSTR>
__device__ __forceinline__ void put_chunk(const u32x4 (&src)[LW], float rn, uint16_t *buf,
                                          float *rn_buf, int wave, int lane) {
#pragma unroll
    for (int t = 0; t < LW; ++t) {
        const int j = wave + 4 * t;
        if (j < INSTR) *reinterpret_cast<u32x4 *>(buf + (j * 64 + lane) * 8) = src[t];
    }
    if (wave == 0 && lane < 16) rn_buf[lane] = rn;
}

// Grids of at most one workgroup per CU (small batches): the scan with
// chunks loaded into registers up to four chunks ahead and written to LDS by
// the waves themselves, so that four chunks stay in flight while one is
// computed (one workgroup per CU has no second workgroup to cover a chunk's
// load).
// The LDS image keeps padded rows (SLOTS_ROW words); the fragments, MFMA
// order and rn - 2 q.r are those of knn_scan_kernel.
// Scan cells (query block, slice) in slice-major order: cell = slice *
// query_blocks + query block. Workgroups are dealt to the 8 XCDs in turn;
// the scan launches 8 * ceil(cells / 8) of them and workgroup b takes cell
// (b mod 8) * ceil(cells / 8) + b / 8, so each XCD runs one contiguous run
// of cells and the query blocks of a slice stream that slice's references
// through the same L2 together. Returns false for the padding workgroups.
// Query rows of a block of kTQ = 128 across the four waves: MFMA row r of
// sub-block qa on wave w is query qa * 64 + 16 w + r, so a batch of fewer
// than 128 rows still gives every wave its share of the survivors. Lane
// l < 32 of wave w owns the thresholds and counts of query own_query(w, l).
__device__ inline int own_query(int wave, int l) { return (l >> 4) * 64 + wave * 16 + (l & 15); }

__device__ inline bool scan_cell(int64_t nq, int slices, int64_t &q_block, int &slice,
                                 int bq = kTQ) {
    const int64_t qblocks = (nq + bq - 1) / bq;
    const int64_t cells = qblocks * slices;
    const int64_t chunk = (cells + 7) / 8;
    const int64_t cell = (blockIdx.x % 8) * chunk + blockIdx.x / 8;
    if (cell >= cells) return false;
    slice = static_cast<int>(cell / qblocks);
    q_block = (cell % qblocks) * bq;
    return true;
}

template <int KSTEPS>
__global__ __launch_bounds__(256, 1) void knn_scan_small_kernel(
    const uint16_t *__restrict__ qb, int64_t nq, const uint16_t *__restrict__ rb,
    const float *__restrict__ rnorm, int64_t nr, int64_t per_slice, int slices,
    float *__restrict__ cand_d, int32_t *__restrict__ cand_i, int32_t *__restrict__ cand_n,
    uint32_t *__restrict__ tau_key, int exclude_self, int64_t query_offset) {
    constexpr int DP = KSTEPS * 32;
    constexpr int SLOTS_ROW = DP / 8 + 1;       // 16-byte words per padded LDS row
    constexpr int ROW = SLOTS_ROW * 8;          // bf16 per padded row
    constexpr int INSTR = (kTR * SLOTS_ROW + 63) / 64;   // wave-wide 1 KB loads per chunk
    constexpr int BUF = INSTR * 64 * 8;         // bf16 per buffer (whole 1 KB loads)
    constexpr int kRefresh = 32;                // chunks between reads of the global thresholds
    __shared__ __attribute__((aligned(16))) uint16_t sB[2 * BUF];
    __shared__ float s_rn[2][16];   // reference norms of each buffered chunk
    __shared__ float s_tau[kTQ];
    __shared__ int s_cnt[kTQ];

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    int64_t q_block = 0;
    int slice = 0;
    if (!scan_cell(nq, slices, q_block, slice)) return;
    const int64_t r_begin = static_cast<int64_t>(slice) * per_slice;
    const int64_t r_end = min<int64_t>(nr, r_begin + per_slice);

    // Chunk loads into registers (one of SETS sets), and from there into an LDS
    // buffer: LDS word `slot` of the padded image holds row slot / SLOTS_ROW,
    // word slot % SLOTS_ROW (the last word of a row and the slots past the
    // chunk are padding); past the slice they load harmless in-bounds words.
    constexpr int LW = (INSTR + 3) / 4;   // staging loads per wave (some use LW - 1)
    // register sets of staged chunks: chunk i (counted from the slice's
    // start) is in set i mod SETS and LDS buffer i mod 2; four sets keep four
    // chunks in flight, two where the query fragments leave no room for more
    constexpr int SETS = KSTEPS >= 32 ? 2 : 4;
    u32x4 pre[SETS][LW];
    float rnv[SETS];

    // A fragments: lane holds row (lane % 16) and K elements 8*(lane/16)..+7
    // of each 32-wide K step, for two 16-query sub-blocks.
    bf16x8 A[2][KSTEPS];
#pragma unroll
    for (int qa = 0; qa < 2; ++qa) {
        const int64_t row = q_block + qa * 64 + wave * 16 + (lane & 15);
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ++ks) {
            if (row < nq) {
                const uint4 raw = *reinterpret_cast<const uint4 *>(qb + row * DP + ks * 32 + 8 * (lane >> 4));
                A[qa][ks] = __builtin_bit_cast(bf16x8, raw);
            } else {
                A[qa][ks] = __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
            }
        }
    }
    const int ql_own = own_query(wave, lane & 31);
    if (lane < 32) {
        // the query's global threshold so far (+inf, or its sample bound)
        const int64_t q = q_block + ql_own;
        s_tau[ql_own] = q < nq ? key_float(tau_key[q]) : INFINITY;
        s_cnt[ql_own] = 0;
    }
    // C layout: lane holds rows 4*(lane/16)+i (queries), column lane%16
    // (reference). A candidate is kept when dt <= tau, tau being the lower
    // of the slice's own 64th best and the query's best 64th best over all
    // slices (both bound the query's final 64th best from above).
    float tau_r[2][4];
#pragma unroll
    for (int qa = 0; qa < 2; ++qa)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t q = q_block + qa * 64 + wave * 16 + 4 * (lane >> 4) + i;
            tau_r[qa][i] = q < nq ? key_float(tau_key[q]) : INFINITY;
        }
    // the query fragments and thresholds have landed before the chunk loads
    // start, so no wait in the loop counts them
    wait_vmcnt<0>();
    if (r_begin < r_end) {
#pragma unroll
        for (int s = 0; s < SETS; ++s)
            fetch_chunk<LW, INSTR, SLOTS_ROW, DP>(pre[s], rnv[s], rb, rnorm, r_begin + s * kTR,
                                                  r_end, wave, lane);
        put_chunk<LW, INSTR>(pre[0], rnv[0], sB, &s_rn[0][0], wave, lane);
        fetch_chunk<LW, INSTR, SLOTS_ROW, DP>(pre[0], rnv[0], rb, rnorm, r_begin + SETS * kTR,
                                              r_end, wave, lane);
    }

    // One chunk: wait for its buffer, compute it against the queries and keep
    // the survivors. The chunk's set is a constant so that each of the SETS
    // unrolled copies stores and refills a fixed register set.
    int since = 0;
    auto scan_chunk = [&](int64_t chunk, auto setc) {
        constexpr int set = decltype(setc)::value;
        const int buf = set & 1;
        // chunk's LDS writes are done and the other buffer is free; the fences
        // order LDS only, so the chunks in flight are not waited on
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        if (++since == kRefresh) {
            // thresholds other slices have published for this wave's queries
            since = 0;
            const int64_t q = q_block + ql_own;
            if (lane < 32 && q < nq) {
                const float g = key_float(tau_key[q]);
                if (g < s_tau[ql_own]) s_tau[ql_own] = g;
            }
#pragma unroll
            for (int qa = 0; qa < 2; ++qa)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    tau_r[qa][i] = s_tau[qa * 64 + wave * 16 + 4 * (lane >> 4) + i];
        }
        const float rn = s_rn[buf][lane & 15];
        // registers: chunks + kTR .. + SETS kTR; chunk + kTR goes to the other
        // LDS buffer and chunk + (SETS + 1) kTR into its set. Done past the
        // slice's end too (harmless words into a buffer nobody reads): a
        // conditional refill would leave the compiler unsure which register
        // set holds the newest loads, and it would then wait for all of them.
        constexpr int next = (set + 1) % SETS;
        put_chunk<LW, INSTR>(pre[next], rnv[next], sB + (buf ^ 1) * BUF, &s_rn[buf ^ 1][0], wave,
                             lane);
        fetch_chunk<LW, INSTR, SLOTS_ROW, DP>(pre[next], rnv[next], rb, rnorm,
                                              chunk + (SETS + 1) * kTR, r_end, wave, lane);
        const uint16_t *B = sB + buf * BUF;
        floatx4 acc[2] = {floatx4{0.0f, 0.0f, 0.0f, 0.0f}, floatx4{0.0f, 0.0f, 0.0f, 0.0f}};
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ++ks) {
            const uint4 raw = *reinterpret_cast<const uint4 *>(&B[(lane & 15) * ROW + ks * 32 + 8 * (lane >> 4)]);
            const bf16x8 bfrag = __builtin_bit_cast(bf16x8, raw);
#pragma unroll
            for (int qa = 0; qa < 2; ++qa)
                acc[qa] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[qa][ks], bfrag, acc[qa], 0, 0, 0);
        }
        const int64_t ref = chunk + (lane & 15);
        bool any = false;
#pragma unroll
        for (int qa = 0; qa < 2; ++qa)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int ql = qa * 64 + wave * 16 + 4 * (lane >> 4) + i;
                const int64_t q = q_block + ql;
                const float dt = rn - 2.0f * acc[qa][i];
                const bool keep = ref < r_end && q < nq && dt <= tau_r[qa][i] &&
                                  !(exclude_self && ref == query_offset + q);
                const uint64_t m = __ballot(keep);
                if (m) {
                    // the 16 lanes of a lane group share query ql; positions
                    // by ballot prefix, count kept in the wave-private s_cnt
                    const int base_n = s_cnt[ql];
                    const uint32_t gm = static_cast<uint32_t>(m >> (16 * (lane >> 4))) & 0xFFFFu;
                    if (keep) {
                        const int pos = base_n + __popc(gm & ((1u << (lane & 15)) - 1u));
                        const int64_t base = (q * slices + slice) * kCap;
                        cand_d[base + pos] = dt;
                        cand_i[base + pos] = static_cast<int32_t>(ref);
                    }
                    if ((lane & 15) == 0) s_cnt[ql] = base_n + __popc(gm);
                    any = true;
                }
            }
        if (any) {
            // this wave's buffers that could overflow with the next chunk
            const int n_l = lane < 32 ? s_cnt[ql_own] : 0;
            uint64_t full = __ballot(lane < 32 && n_l > kCap - kTR);
            if (full) {
                __threadfence_block();   // the wave's appends are visible to all its lanes
                while (full) {
                    const int l = __builtin_ctzll(full);
                    full &= full - 1;
                    const int ql = own_query(wave, l);
                    const int64_t q = q_block + ql;
                    const int n = __shfl(n_l, l, 64);
                    const int64_t base = (q * slices + slice) * kCap;
                    const float tau = compact_buffer(cand_d + base, cand_i + base, n, lane);
                    if (lane == 0) {
                        // publish; the query's bound is the lower of both
                        const uint32_t old = atomicMin(&tau_key[q], float_key(tau));
                        const float g = key_float(old);
                        s_tau[ql] = g < tau ? g : tau;
                        s_cnt[ql] = kTop;
                    }
                }
#pragma unroll
                for (int qa = 0; qa < 2; ++qa)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        tau_r[qa][i] = s_tau[qa * 64 + wave * 16 + 4 * (lane >> 4) + i];
            }
        }
    };
    // groups of SETS chunks, then up to SETS - 1 more
    int64_t chunk = r_begin;
    for (; chunk + (SETS - 1) * kTR < r_end; chunk += SETS * kTR) {
        scan_chunk(chunk, std::integral_constant<int, 0>{});
        scan_chunk(chunk + kTR, std::integral_constant<int, 1>{});
        if constexpr (SETS == 4) {
            scan_chunk(chunk + 2 * kTR, std::integral_constant<int, 2>{});
            scan_chunk(chunk + 3 * kTR, std::integral_constant<int, 3>{});
        }
    }
    if (chunk < r_end) scan_chunk(chunk, std::integral_constant<int, 0>{});
    if constexpr (SETS == 4) {
        if (chunk + kTR < r_end) scan_chunk(chunk + kTR, std::integral_constant<int, 1>{});
        if (chunk + 2 * kTR < r_end) scan_chunk(chunk + 2 * kTR, std::integral_constant<int, 2>{});
    }
    __threadfence_block();
    __syncthreads();
    finish_cell(s_cnt, q_block, nq, slices, slice, cand_d, cand_i, cand_n, wave, lane);
}

// The prefilter scan of one (query block, slice) for grids of two workgroups
// per CU. Each chunk of 16 references is staged with global_load_lds into
// the other half of an LDS double buffer while the current one feeds the
// MFMAs.
//
// LDS image of a chunk, K-step major: KSTEPS blocks of 1 KB, block ks holds
// the 16 rows x 32 features of MFMA K step ks, row r's four 16-byte words at
// byte 64 r + 16 (w ^ ((r >> 2) & 3)) (the XOR puts any 16 rows of one word
// on distinct banks). One DMA instruction fills one block: lane l loads row
// l >> 2, word (l & 3) ^ ((l >> 4) & 3), so a lane's source offset is the
// same for every block and chunk up to a multiple of 64 bytes (no address
// arithmetic per 16-byte word), and a B fragment is one ds_read_b128 at a
// lane-constant offset plus 1 KB per K step.
//
// Registers: the A fragments take 8 KSTEPS VGPRs (200 at 784 features) of
// the 256 that two waves per SIMD allow. The rest must hold a ring of RING
// B fragments (the read of K step ks + RING is in flight while step ks feeds
// the MFMAs), the accumulators and the epilogue: thresholds are read from
// LDS per chunk and the epilogue's indices are 32-bit, so nothing spills. A
// spill would be costly: with an LDS DMA in flight the compiler turns every
// wait on a vector memory load (a scratch reload included) into vmcnt(0).
// Above 25 K steps the A fragments alone fill 256 VGPRs, so those widths
// run one wave per SIMD (512 VGPRs).
//
// SAMPLE: the sample distances instead. The rows are the sample entries
// 0 .. sample_m - 1, in blocks of 16: entry i is reference
// (i / 16) nr / (sample_m / 16) + i mod 16, so a chunk is 16 consecutive
// rows as in the scan. They are sliced like the references, and every
// (query, entry) value is written to sample_dist[q * sample_m + i] (+inf
// for an excluded self pair), so each equals the value the scan computes
// for the same (query, reference).
//
// EXCL: the self pairs may need excluding (exclude_self); without it the
// keep test is one compare per (query, reference).
//
// QA: 16-query sub-blocks per wave, so a block holds 64 QA queries. Batches
// of at most 64 rows take QA = 1: half the A fragments and half the MFMAs
// per chunk, and three workgroups fit a CU.
template <int KSTEPS, int RING, bool SAMPLE, bool EXCL, int QA>
__global__ __launch_bounds__(256, KSTEPS > 25 ? 1 : 2) void knn_scan_kernel(
    const uint16_t *__restrict__ qb, int64_t nq, const uint16_t *__restrict__ rb,
    const float *__restrict__ rnorm, int64_t nr, int64_t per_slice, int slices,
    float *__restrict__ cand_d, int32_t *__restrict__ cand_i, int32_t *__restrict__ cand_n,
    uint32_t *__restrict__ tau_key, int exclude_self, int64_t query_offset, int64_t sample_m,
    float *__restrict__ sample_dist) {
    constexpr int DP = KSTEPS * 32;
    constexpr int BUF = KSTEPS * 512;   // bf16 per buffer
    constexpr int kRefresh = 32;
    constexpr int BQ = 64 * QA;         // queries per block
    static_assert(QA == 1 || QA == 2, "one or two sub-blocks per wave");
    __shared__ __attribute__((aligned(16))) uint16_t sB[2 * BUF];
    __shared__ __attribute__((aligned(16))) float s_tau[kTQ];
    __shared__ int s_cnt[kTQ];

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    __builtin_assume(wave >= 0 && wave < 4);   // lost by readfirstlane; simplifies the DMA predicates
    int64_t q_block64 = 0;
    int slice = 0;
    if (!scan_cell(nq, slices, q_block64, slice, BQ)) return;
    const int q_block = static_cast<int>(q_block64);
    const int inr = static_cast<int>(SAMPLE ? sample_m : nr);   // rows: references or entries
    const int r_begin = static_cast<int>(slice * per_slice);
    const int r_end = static_cast<int>(min<int64_t>(inr, r_begin + per_slice));
    const int qnq = static_cast<int>(nq);
    // the reference of row r
    auto ref_of = [&](int r) -> int {
        return SAMPLE ? static_cast<int>(static_cast<int64_t>(r / kTR) * nr / (sample_m / kTR)) + r % kTR
                      : r;
    };

    // DMA lane mapping (row and source word of this lane's 16 bytes)
    const int d_row = lane >> 2;
    const int d_word = (lane & 3) ^ ((d_row >> 2) & 3);
    const uint16_t *const d_src = rb + static_cast<int64_t>(d_row) * DP + d_word * 8;
    auto stage = [&](int buf, int chunk) {
        // rows past the last (last chunk only) load row 0 instead
        const uint16_t *src;
        if constexpr (SAMPLE) {
            const int r = chunk + d_row < inr ? chunk + d_row : 0;
            src = rb + static_cast<int64_t>(ref_of(r)) * DP + d_word * 8 + wave * 32;
        } else {
            src = d_src + static_cast<int64_t>(chunk) * DP + wave * 32;
            if (chunk + kTR > inr && chunk + d_row >= inr) src = rb + d_word * 8 + wave * 32;
        }
#pragma unroll
        for (int t = 0; t < (KSTEPS + 3) / 4; ++t) {
            if (wave + 4 * t < KSTEPS)
                __builtin_amdgcn_global_load_lds(
                    (const void __attribute__((address_space(1))) *)(src + t * 128),
                    (void __attribute__((address_space(3))) *)(sB + buf * BUF + (wave + 4 * t) * 512), 16, 0, 0);
        }
    };
    if (r_begin < r_end) stage(0, r_begin);

    bf16x8 A[QA][KSTEPS];
#pragma unroll
    for (int qa = 0; qa < QA; ++qa) {
        const int64_t row = q_block + qa * 64 + wave * 16 + (lane & 15);
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ++ks) {
            if (row < nq) {
                const uint4 raw = *reinterpret_cast<const uint4 *>(qb + row * DP + ks * 32 + 8 * (lane >> 4));
                A[qa][ks] = __builtin_bit_cast(bf16x8, raw);
            } else {
                A[qa][ks] = __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
            }
        }
    }
    const int ql_own = own_query(wave, lane & 31);
    constexpr int kOwners = 16 * QA;    // lanes that own a query's threshold and count
    if (!SAMPLE && lane < kOwners) {
        // queries past nq get -inf: no distance passes dt <= -inf
        const int q = q_block + ql_own;
        s_tau[ql_own] = q < qnq ? key_float(tau_key[q]) : -INFINITY;
        s_cnt[ql_own] = 0;
    }
    // a row past the slice gets a NaN norm: its distances fail every keep test
    const float kNoRow = __builtin_nanf("");
    float rn_next = kNoRow;
    if (r_begin + (lane & 15) < r_end) rn_next = rnorm[ref_of(r_begin + (lane & 15))];
    // lane's C rows: queries ql_lane + qa * 64 + i
    const int ql_lane = wave * 16 + 4 * (lane >> 4);
    const int self_base = static_cast<int>(query_offset) + q_block + ql_lane;
    // B fragment: row lane & 15, word lane >> 4 of each K block
    const int b_row = lane & 15;
    const uint16_t *const Bl0 = sB + b_row * 32 + (((lane >> 4) ^ ((b_row >> 2) & 3)) * 8);

    // SAMPLE: a chunk's values are written in the next chunk, after its DMA
    // is issued, so that the wait at the top of a chunk (vmcnt(0)) finds the
    // stores as far along as the DMA instead of just issued
    float pend[QA][4] = {};
    int pend_chunk = -1;
    auto write_pending = [&]() {
        const int e = pend_chunk + (lane & 15);
        if (pend_chunk >= 0 && e < r_end) {
            float *out = sample_dist + static_cast<int64_t>(q_block + ql_lane) * sample_m + e;
#pragma unroll
            for (int qa = 0; qa < QA; ++qa)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (q_block + ql_lane + qa * 64 + i < qnq)
                        out[static_cast<int64_t>(qa * 64 + i) * sample_m] = pend[qa][i];
        }
    };

    int since = 0, buf = 0;
    for (int chunk = r_begin; chunk < r_end; chunk += kTR, buf ^= 1) {
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): this lane's loads of the current buffer
        __syncthreads();                      // every lane's loads landed; the other buffer is free
        if (!SAMPLE && ++since == kRefresh) {
            since = 0;
            const int q = q_block + ql_own;
            if (lane < kOwners && q < qnq) {
                const float g = key_float(tau_key[q]);
                if (g < s_tau[ql_own]) s_tau[ql_own] = g;
            }
        }
        const float rn = rn_next;
        if (chunk + kTR < r_end) {
            stage(buf ^ 1, chunk + kTR);
            const int nref = chunk + kTR + (lane & 15);
            rn_next = nref < r_end ? rnorm[ref_of(nref)] : kNoRow;
        }
        if constexpr (SAMPLE) write_pending();
        const uint16_t *Bl = Bl0 + buf * BUF;
        floatx4 acc[QA];
#pragma unroll
        for (int qa = 0; qa < QA; ++qa) acc[qa] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        uint4 ring[RING];
#pragma unroll
        for (int r = 0; r < RING; ++r)
            if (r < KSTEPS) ring[r] = *reinterpret_cast<const uint4 *>(Bl + r * 512);
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ++ks) {
            const bf16x8 bfrag = __builtin_bit_cast(bf16x8, ring[ks % RING]);
            if (ks + RING < KSTEPS) ring[ks % RING] = *reinterpret_cast<const uint4 *>(Bl + (ks + RING) * 512);
#pragma unroll
            for (int qa = 0; qa < QA; ++qa)
                acc[qa] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[qa][ks], bfrag, acc[qa], 0, 0, 0);
        }
        // issue order: the RING first reads, then per K step its two MFMAs
        // and the read RING steps ahead (the scheduler would otherwise batch
        // the reads and wait for each batch)
#pragma unroll
        for (int r = 0; r < RING; ++r) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ++ks) {
            __builtin_amdgcn_sched_group_barrier(0x008, QA, 0);
            if (ks + RING < KSTEPS) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
        if constexpr (SAMPLE) {
            // the entry's own query, counted from the lane's first query
            const int64_t self0 = ref_of(min(chunk + (lane & 15), r_end - 1)) - query_offset -
                                  (q_block + ql_lane);
#pragma unroll
            for (int qa = 0; qa < QA; ++qa)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float dt = rn - 2.0f * acc[qa][i];
                    pend[qa][i] = exclude_self && self0 == qa * 64 + i ? INFINITY : dt;
                }
            pend_chunk = chunk;
            continue;
        }
        const int ref = chunk + (lane & 15);
        const int self_d = EXCL && exclude_self ? ref - self_base : -1;
        bool any = false;
#pragma unroll
        for (int qa = 0; qa < QA; ++qa) {
            const floatx4 tau = *reinterpret_cast<const floatx4 *>(&s_tau[ql_lane + qa * 64]);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int ql = ql_lane + qa * 64 + i;
                const float dt = rn - 2.0f * acc[qa][i];
                const bool keep = dt <= tau[i] && (!EXCL || self_d != qa * 64 + i);
                const uint64_t m = __builtin_amdgcn_ballot_w64(keep);
                if (m) {
                    const int base_n = s_cnt[ql];
                    const uint32_t gm = static_cast<uint32_t>(m >> (16 * (lane >> 4))) & 0xFFFFu;
                    if (keep) {
                        const int pos = base_n + __popc(gm & ((1u << (lane & 15)) - 1u));
                        int qv = q_block + ql;
                        asm volatile("" : "+v"(qv));   // keeps the address computation in the branch
                        const int64_t base = (static_cast<int64_t>(qv) * slices + slice) * kCap;
                        cand_d[base + pos] = dt;
                        cand_i[base + pos] = ref;
                    }
                    if ((lane & 15) == 0) s_cnt[ql] = base_n + __popc(gm);
                    any = true;
                }
            }
        }
        if (any) {
            const int n_l = lane < kOwners ? s_cnt[ql_own] : 0;
            uint64_t full = __builtin_amdgcn_ballot_w64(lane < kOwners && n_l > kCap - kTR);
            if (full) {
                __threadfence_block();
                while (full) {
                    const int l = __builtin_ctzll(full);
                    full &= full - 1;
                    const int ql = own_query(wave, l);
                    const int q = q_block + ql;
                    const int n = __shfl(n_l, l, 64);
                    const int64_t base = (static_cast<int64_t>(q) * slices + slice) * kCap;
                    const float tau = compact_buffer(cand_d + base, cand_i + base, n, lane);
                    if (lane == 0) {
                        const uint32_t old = atomicMin(&tau_key[q], float_key(tau));
                        const float g = key_float(old);
                        s_tau[ql] = g < tau ? g : tau;
                        s_cnt[ql] = kTop;
                    }
                }
                __threadfence_block();
            }
        }
    }
    if constexpr (SAMPLE) {
        write_pending();
    } else {
        __threadfence_block();
        __syncthreads();
        finish_cell(s_cnt, q_block, nq, slices, slice, cand_d, cand_i, cand_n, wave, lane, BQ);
    }
}

// Workgroup q: the kTop-th smallest of the query's m (at most 64 WAVES
// KPT) sample distances by a most-significant-byte-first radix select over
// their order-preserving keys, widened by 2^-20 of its magnitude, becomes
// the query's global threshold (+inf when fewer than kTop sample distances
// are finite). The keys are read once into registers (KPT per thread);
// each wave counts into its own histogram. The first (exponent) byte takes
// only a few values, so there a wave adds each distinct byte once (lanes
// grouped by ballot) instead of 64 atomics on one bin.
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void knn_sample_select_kernel(const float *__restrict__ dist,
                                                                       int m,
                                                                       uint32_t *__restrict__ tau_key) {
    constexpr int T = 64 * WAVES;
    constexpr int KPT = kSelectKeys;
    static_assert(WAVES >= 4, "threads 0..255 own the 256 bins");
    __shared__ uint32_t s_hist[WAVES][256];
    __shared__ uint32_t s_wsum[4];
    __shared__ uint32_t s_sel[2];   // chosen byte, rank left within it
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t q = blockIdx.x;
    const float *row = dist + q * m;
    uint32_t key[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        const int i = j * T + tid;
        key[j] = i < m ? float_key(row[i]) : 0u;
    }
    uint32_t prefix = 0u, mask = 0u, want = kTop - 1;   // 0-based rank of the kTop-th smallest
    // Three byte passes fix the top 24 bits of the kTop-th smallest key; the
    // bound is the largest value with those bits (at most 2^-15 relative
    // above it), still an upper bound of every query's final kTop-th best.
    for (int shift = 24; shift >= 8; shift -= 8) {
        for (int e = tid; e < WAVES * 256; e += T) (&s_hist[0][0])[e] = 0u;
        __syncthreads();
        if (shift == 24) {
#pragma unroll
            for (int j = 0; j < KPT; ++j) {
                const uint32_t b = key[j] >> 24;
                uint64_t todo = __builtin_amdgcn_ballot_w64(j * T + tid < m);
                while (todo) {
                    const uint32_t lb = __builtin_amdgcn_readlane(b, static_cast<int>(__builtin_ctzll(todo)));
                    const uint64_t same = todo & __builtin_amdgcn_ballot_w64(b == lb);
                    if (lane == static_cast<int>(__builtin_ctzll(todo)))
                        atomicAdd(&s_hist[wave][lb], static_cast<uint32_t>(__popcll(same)));
                    todo &= ~same;
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < KPT; ++j)
                if (j * T + tid < m && (key[j] & mask) == prefix)
                    atomicAdd(&s_hist[wave][(key[j] >> shift) & 255u], 1u);
        }
        __syncthreads();
        // the byte whose count range [below, below + count) holds rank want
        // (threads 0..255, one bin each)
        uint32_t c = 0u, incl = 0u;
        if (tid < 256) {
#pragma unroll
            for (int w = 0; w < WAVES; ++w) c += s_hist[w][tid];
            incl = static_cast<uint32_t>(wave_inclusive_sum(static_cast<int>(c)));
            if (lane == 63) s_wsum[wave] = incl;
        }
        __syncthreads();
        if (tid < 256) {
            for (int w = 0; w < wave; ++w) incl += s_wsum[w];
            const uint32_t below = incl - c;
            if (c > 0u && below <= want && want < incl) {
                s_sel[0] = static_cast<uint32_t>(tid);
                s_sel[1] = want - below;
            }
        }
        __syncthreads();
        prefix |= s_sel[0] << shift;
        mask |= 255u << shift;
        want = s_sel[1];
        __syncthreads();
    }
    if (tid == 0) {
        const float kth = key_float(prefix | 0xFFu);
        // sample values equal the scan's; the widening only guards the bound
        const float bound = m >= kTop && kth < INFINITY ? kth + fabsf(kth) * 0x1.0p-20f : INFINITY;
        tau_key[q] = float_key(bound);
    }
}

// Merge the sorted per-slice lists of each query into its best kTop
// (sorted by distance, then id). One workgroup of WAVES waves per query:
// the query's listed slices are dealt to the waves in turn; in a wave, lane
// l holds entry l of the running list A, and each slice list B is loaded
// reversed, so A followed by B is bitonic: a half-cleaner keeps the 64
// smallest and a bitonic merge sorts them, all in registers. The wave lists
// then merge pairwise in log2(WAVES) rounds. The best kTop under the total
// order (distance, id) does not depend on the order of the merges. Every
// member of the query's best 64 is at most its published threshold, so
// slices whose best candidate is above it are not listed. A wave keeps
// kMergeAhead slices in flight; a slice's 64 entries are read whatever its
// count (words of its own buffer) and masked when merged, so a fetch is one
// round trip.
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void knn_merge_kernel(const float *cand_d,
                                                               const int32_t *cand_i,
                                                               const int32_t *cand_n,
                                                               const uint32_t *tau_key, int64_t nq,
                                                               int slices, float *top_d,
                                                               int32_t *top_i, int32_t *top_n,
                                                               int32_t *open_count) {
    static_assert((WAVES & (WAVES - 1)) == 0, "a power of two of waves");
    constexpr int PF = kMergeAhead;
    __shared__ int32_t s_list[kMaxSlices];
    __shared__ int s_count;
    __shared__ float s_d[WAVES][kTop];
    __shared__ int32_t s_i[WAVES][kTop];
    __shared__ int s_na[WAVES];
    const int64_t q = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (q == 0 && tid == 0) *open_count = 0;   // the rerank counts the uncertified queries
    if (q >= nq) return;
    const float tau = key_float(tau_key[q]);
    if (tid == 0) s_count = 0;
    __syncthreads();
    for (int sl = tid; sl < slices; sl += 64 * WAVES) {
        // list order is free: the merged result does not depend on it
        const int n = cand_n[q * slices + sl];
        if (n > 0 && cand_d[(q * slices + sl) * kCap] <= tau) s_list[atomicAdd(&s_count, 1)] = sl;
    }
    __syncthreads();
    const int listed = s_count;
    float ad = INFINITY;
    int32_t ai = INT32_MAX;
    int na = 0;
    // merges sorted list B (entry 63 - lane in this lane) into the wave's list
    auto merge_in = [&](float bd, int32_t bi) {
        if (bd < ad || (bd == ad && bi < ai)) {
            ad = bd;
            ai = bi;
        }
#pragma unroll
        for (int j = 32; j > 0; j >>= 1) {
            const float dp = __shfl_xor(ad, j, 64);
            const int32_t ip = __shfl_xor(ai, j, 64);
            const bool mine_gt = ad > dp || (ad == dp && ai > ip);
            const bool lower = (lane & j) == 0;
            if (lower == mine_gt) {
                ad = dp;
                ai = ip;
            }
        }
    };
    const int r = kTop - 1 - lane;   // this lane's entry of a reversed list
    float pd[PF];
    int32_t pi[PF];
    int pn[PF];
    auto fetch = [&](int li, int p) {
        const int64_t cell = q * slices + s_list[li];
        pn[p] = cand_n[cell];
        pd[p] = cand_d[cell * kCap + r];
        pi[p] = cand_i[cell * kCap + r];
    };
#pragma unroll
    for (int p = 0; p < PF; ++p)
        if (wave + p * WAVES < listed) fetch(wave + p * WAVES, p);
    for (int li = wave; li < listed; li += PF * WAVES) {
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            const int lp = li + p * WAVES;
            if (lp < listed) {
                const int nb = pn[p];
                const float bd = r < nb ? pd[p] : INFINITY;
                const int32_t bi = r < nb ? pi[p] : INT32_MAX;
                if (lp + PF * WAVES < listed) fetch(lp + PF * WAVES, p);
                merge_in(bd, bi);
                na = min(kTop, na + nb);
            }
        }
    }
    // pairwise rounds: wave w < h takes in the list of wave w + h
#pragma unroll
    for (int h = WAVES / 2; h >= 1; h /= 2) {
        if (wave >= h && wave < 2 * h) {
            s_d[wave][lane] = ad;
            s_i[wave][lane] = ai;
            if (lane == 0) s_na[wave] = na;
        }
        __syncthreads();
        if (wave < h) {
            merge_in(s_d[wave + h][kTop - 1 - lane], s_i[wave + h][kTop - 1 - lane]);
            na = min(kTop, na + s_na[wave + h]);
        }
    }
    if (wave != 0) return;
    top_d[q * kTop + lane] = ad;
    top_i[q * kTop + lane] = ai;
    if (lane == 0) top_n[q] = na;
}

// Certificate of one query's reranked top k.
// g bounds the relative error of an f32 sum of up to dp products or
// squares, doubled (2 dp u / (1 - dp u), u = 2^-24). With x~ / y~ the BF16
// rows (x - s) / (y - s), every reference left out of the 64 has an
// approximate distance dt >= T, the 64th. |dt - (|y~|^2 - 2 x~.y~)| <=
// c (|y~|^2 + 2 |x~| |y~|) with c = g + 2u, and |y~| <= |x~| + t for
// t = |x~ - y~|, so t >= t_lo, the positive root of
// (1 + c) t^2 + 4 c X t = X2l + T - 3 c X^2 (X, X2l: upper and lower bounds
// of |x~| and |x~|^2). The rounding of the rows is at most err(x) and
// eps |y~| <= eps (X + t), so |x - y| >= t_lo (1 - eps) - err(x) - eps X = L,
// and the reference's f32 distance of y is at least L^2 (1 - g). When that
// exceeds the k-th reranked distance, no left-out reference can enter the
// top k (ties included). The rerank appends uncertified queries to
// fail_list.
__device__ inline bool certified(int n, float T_f, float kth_f, float qn_f, float qerr_f,
                                 uint32_t eps_bits, double g, int force) {
    if (force) return false;
    if (n < kTop) return true;   // every eligible reference was reranked
    const double u = 0x1.0p-24, c = g + 2.0 * u;
    const double T = T_f;
    const double qn = qn_f;
    const double x2l = qn * (1.0 - g), xu = sqrt(qn * (1.0 + g));
    const double K = x2l + T - 3.0 * c * xu * xu;
    double tlo = 0.0;
    if (K > 0.0)
        tlo = (sqrt(16.0 * c * c * xu * xu + 4.0 * (1.0 + c) * K) - 4.0 * c * xu) / (2.0 * (1.0 + c));
    const double eps = __uint_as_float(eps_bits);
    const double L = tlo * (1.0 - eps) - static_cast<double>(qerr_f) - eps * xu;
    const double kth = kth_f;
    return eps < 1.0 && L > 0.0 && L * L * (1.0 - g) * (1.0 - 0x1.0p-40) > kth;
}

// Everything the rerank needs besides the rows: the certificate's inputs
// and the list of queries it leaves open.
struct CertifyArgs {
    const float *top_d;
    const float *qnorm;
    const float *qerr;
    const uint32_t *eps_key;
    double g;
    int force;
    int32_t *fail_list;
    int32_t *fail_count;
};

// The reranked k of one query from wave 0's lanes (lane c: candidate c, its
// exact squared distance `acc`): ranks by (distance, id), the k best in
// order with their square-rooted distances, the k-th squared distance for
// the exact rescan, and sentinels (id nr, distance sqrt(FLT_MAX)) past n.
// Then the query's certificate; an uncertified query is appended to
// fail_list.
__device__ inline void rerank_output(float acc, int32_t id, int n, int lane, int64_t q, int k,
                                     int64_t nr, int32_t *out_i, float *out_d, float *kth_d2,
                                     const CertifyArgs &ca) {
    const float dist = lane < n ? acc : INFINITY;
    int rank = 0;
#pragma unroll 8
    for (int f = 0; f < 64; ++f) {
        const float df = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dist), f));
        const int32_t idf = __builtin_amdgcn_readlane(id, f);
        rank += (df < dist || (df == dist && idf < id)) ? 1 : 0;
    }
    if (lane < n && rank < k) {
        out_i[q * k + rank] = id;
        out_d[q * k + rank] = sqrtf(dist);
        if (rank == k - 1) kth_d2[q] = dist;
    }
    if (lane == 0 && n < k) kth_d2[q] = INFINITY;
    if (lane >= n && lane < k) {
        out_i[q * k + lane] = static_cast<int32_t>(nr);
        out_d[q * k + lane] = sqrtf(FLT_MAX);
    }
    const uint64_t at_k = __builtin_amdgcn_ballot_w64(lane < n && rank == k - 1);
    const float kth = at_k ? __int_as_float(__builtin_amdgcn_readlane(
                                 __float_as_int(dist), static_cast<int>(__builtin_ctzll(at_k))))
                           : INFINITY;
    if (lane == 0 && !certified(n, ca.top_d[q * kTop + kTop - 1], kth, ca.qnorm[q], ca.qerr[q],
                                *ca.eps_key, ca.g, ca.force))
        ca.fail_list[atomicAdd(ca.fail_count, 1)] = static_cast<int32_t>(q);
}

// Exact rerank: one workgroup of four waves per query, one lane of wave 0
// per candidate; the distance is the reference's sequential f32 sum over the
// features (no contraction in this build). The four waves load a
// 128-feature block of all 64 candidate rows at once (32 loads in flight
// per lane, every load part of a contiguous row segment) and put the
// squared differences (q - r)^2, each rounded to f32 as the sum's own
// product is, into LDS; wave 0 then only adds the block in feature order.
__global__ __launch_bounds__(256) void knn_rerank_kernel(const float *query, int64_t nq,
                                                              const float *ref, int64_t nr,
                                                              int64_t d, const int32_t *top_i,
                                                              const int32_t *top_n, int k,
                                                              int32_t *out_i, float *out_d,
                                                              float *kth_d2, CertifyArgs ca) {
    constexpr int FB = 128;
    __shared__ float s_r[kTop][FB + 1];
    const int64_t q = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (q >= nq) return;
    const int n = top_n[q];
    const int32_t id = lane < n ? top_i[q * kTop + lane] : INT32_MAX;
    float acc = 0.0f;
    for (int64_t f0 = 0; f0 < d; f0 += FB) {
        const int fb = static_cast<int>(min<int64_t>(FB, d - f0));
        // wave w: candidates 16w .. 16w + 15, features lane and lane + 64
        const float q0 = lane < fb ? query[q * d + f0 + lane] : 0.0f;
        const float q1 = lane + 64 < fb ? query[q * d + f0 + 64 + lane] : 0.0f;
        float v[16][2];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int c = 16 * wave + u;
            const int64_t row = __shfl(id, c, 64);
            v[u][0] = (c < n && lane < fb) ? ref[row * d + f0 + lane] : 0.0f;
            v[u][1] = (c < n && lane + 64 < fb) ? ref[row * d + f0 + 64 + lane] : 0.0f;
        }
        __syncthreads();   // wave 0 is done with the previous block
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int c = 16 * wave + u;
            if (c < n) {
                const float e0 = q0 - v[u][0], e1 = q1 - v[u][1];
                s_r[c][lane] = e0 * e0;
                s_r[c][lane + 64] = e1 * e1;
            }
        }
        __syncthreads();
        if (wave == 0 && lane < n) {
#pragma unroll 8
            for (int f = 0; f < fb; ++f) acc += s_r[lane][f];
        }
    }
    if (wave != 0) return;
    rerank_output(acc, id, n, lane, q, k, nr, out_i, out_d, kth_d2, ca);
}

// knn_rerank_kernel for small batches with d <= 64 kRerankSlots: sixteen
// waves load all 64 candidate rows at once (wave w: candidates 4w .. 4w + 3,
// lane l: features l + 64 t), so the rows arrive in one round trip instead
// of one per 128-feature block. They pass through LDS in 512-feature
// phases as squared differences, and wave 0 adds each candidate's features
// in order with one accumulator, exactly the operations of
// knn_rerank_kernel.
__global__ __launch_bounds__(1024) void knn_rerank_wide_kernel(const float *query, int64_t nq,
                                                               const float *ref, int64_t nr,
                                                               int64_t d, const int32_t *top_i,
                                                               const int32_t *top_n, int k,
                                                               int32_t *out_i, float *out_d,
                                                               float *kth_d2, CertifyArgs ca) {
    constexpr int PH = 512;                 // features per LDS phase
    constexpr int TPP = PH / 64;            // register slots per phase
    constexpr int PHASES = kRerankSlots / TPP;
    __shared__ float s_r[kTop][PH + 1];
    const int64_t q = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (q >= nq) return;
    const int n = top_n[q];
    const int32_t id = lane < n ? top_i[q * kTop + lane] : INT32_MAX;
    float qv[kRerankSlots];
#pragma unroll
    for (int t = 0; t < kRerankSlots; ++t) {
        const int64_t f = lane + 64 * t;
        qv[t] = f < d ? query[q * d + f] : 0.0f;
    }
    float v[4][kRerankSlots];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int c = 4 * wave + u;
        const int64_t row = __shfl(id, c, 64);
#pragma unroll
        for (int t = 0; t < kRerankSlots; ++t) {
            const int64_t f = lane + 64 * t;
            v[u][t] = (c < n && f < d) ? ref[row * d + f] : 0.0f;
        }
    }
    float acc = 0.0f;
#pragma unroll
    for (int ph = 0; ph < PHASES; ++ph) {
        const int64_t f0 = static_cast<int64_t>(ph) * PH;
        if (f0 >= d) break;
        const int fb = static_cast<int>(min<int64_t>(PH, d - f0));
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = 4 * wave + u;
            if (c < n) {
#pragma unroll
                for (int t = 0; t < TPP; ++t) {
                    const float e = qv[ph * TPP + t] - v[u][ph * TPP + t];
                    s_r[c][t * 64 + lane] = e * e;
                }
            }
        }
        __syncthreads();
        if (wave == 0 && lane < n) {
#pragma unroll 8
            for (int f = 0; f < fb; ++f) acc += s_r[lane][f];
        }
    }
    if (wave != 0) return;
    rerank_output(acc, id, n, lane, q, k, nr, out_i, out_d, kth_d2, ca);
}


// Exact rescan of the uncertified queries fail_list[base + slot] for slots
// of this round (kExactQB per workgroup; blockIdx.y = a slice of the
// references). Each lane owns one reference of a 64-row tile and adds the
// squared f32 differences in feature order, the reference's sequential
// distance (feature blocks staged through LDS). Candidates at or below the
// query's reranked k-th distance, an upper bound on its exact k-th, are
// kept by (distance, id); the slice's best 64 go to ex_d / ex_i
// [(slot * slices + slice) * 64] with their count in ex_n.
__global__ __launch_bounds__(64) void knn_exact_scan_kernel(
    const float *__restrict__ query, const float *__restrict__ ref, int64_t nr, int64_t d,
    const int32_t *fail_list, const int32_t *fail_count, int64_t base, int64_t per_slice,
    int slices, const float *kth_d2, int exclude_self, int64_t query_offset, float *ex_d,
    int32_t *ex_i, int32_t *ex_n) {
    constexpr int FB = 64;
    __shared__ float s_q[kExactQB][FB];
    __shared__ float s_r[64][FB + 1];
    __shared__ float s_cd[kExactQB][kCap];
    __shared__ int32_t s_ci[kExactQB][kCap];
    const int lane = threadIdx.x;
    const int64_t f0 = base + static_cast<int64_t>(blockIdx.x) * kExactQB;
    const int64_t count = *fail_count;
    if (f0 >= count) return;
    const int slice = blockIdx.y;
    const int64_t r_begin = static_cast<int64_t>(slice) * per_slice;
    const int64_t r_end = min<int64_t>(nr, r_begin + per_slice);
    int64_t qid[kExactQB];
    float tau[kExactQB];
    int cnt[kExactQB];
#pragma unroll
    for (int b = 0; b < kExactQB; ++b) {
        qid[b] = f0 + b < count ? fail_list[f0 + b] : -1;
        tau[b] = qid[b] >= 0 ? kth_d2[qid[b]] : -INFINITY;
        cnt[b] = 0;
    }
    for (int64_t t0 = r_begin; t0 < r_end; t0 += 64) {
        float acc[kExactQB];
#pragma unroll
        for (int b = 0; b < kExactQB; ++b) acc[b] = 0.0f;
        for (int64_t c0 = 0; c0 < d; c0 += FB) {
            const int fb = static_cast<int>(min<int64_t>(FB, d - c0));
            __syncthreads();
            for (int rr = 0; rr < 64; ++rr) {
                const int64_t row = t0 + rr;
                s_r[rr][lane] = (row < r_end && lane < fb) ? ref[row * d + c0 + lane] : 0.0f;
            }
#pragma unroll
            for (int b = 0; b < kExactQB; ++b)
                s_q[b][lane] = (qid[b] >= 0 && lane < fb) ? query[qid[b] * d + c0 + lane] : 0.0f;
            __syncthreads();
            for (int f = 0; f < fb; ++f) {
                const float rv = s_r[lane][f];
#pragma unroll
                for (int b = 0; b < kExactQB; ++b) {
                    const float delta = s_q[b][f] - rv;
                    acc[b] += delta * delta;
                }
            }
        }
        const int64_t mine = t0 + lane;
#pragma unroll
        for (int b = 0; b < kExactQB; ++b) {
            const bool keep = qid[b] >= 0 && mine < r_end && acc[b] <= tau[b] &&
                              !(exclude_self && mine == query_offset + qid[b]);
            const uint64_t m = __ballot(keep);
            if (!m) continue;
            if (keep) {
                const int pos = cnt[b] + __popcll(m & ((1ull << lane) - 1ull));
                s_cd[b][pos] = acc[b];
                s_ci[b][pos] = static_cast<int32_t>(mine);
            }
            cnt[b] += __popcll(m);
            if (cnt[b] > kCap - 64) {
                // keep the best 64 so the next tile fits
                __syncthreads();
                const float t = compact_buffer(s_cd[b], s_ci[b], cnt[b], lane);
                __syncthreads();
                cnt[b] = kTop;
                tau[b] = fminf(tau[b], t);
            }
        }
    }
#pragma unroll
    for (int b = 0; b < kExactQB; ++b) {
        if (qid[b] < 0) continue;
        const int n = cnt[b];
        __syncthreads();
        if (n > 0) compact_buffer(s_cd[b], s_ci[b], n, lane);
        __syncthreads();
        const int64_t cell = (f0 + b - base) * slices + slice;
        const int kept = n < kTop ? n : kTop;
        if (lane < kept) {
            ex_d[cell * kTop + lane] = s_cd[b][lane];
            ex_i[cell * kTop + lane] = s_ci[b][lane];
        }
        if (lane == 0) ex_n[cell] = kept;
    }
}

// Merges sorted list B (entry 63 - lane in this lane) into the sorted list
// A held one entry per lane (half-cleaner of the bitonic A ++ reverse(B),
// then a bitonic merge), keeping the best 64 by (distance, id).
__device__ inline void merge_reversed(float &ad, int32_t &ai, float bd, int32_t bi, int lane) {
    if (bd < ad || (bd == ad && bi < ai)) {
        ad = bd;
        ai = bi;
    }
#pragma unroll
    for (int j = 32; j > 0; j >>= 1) {
        const float dp = __shfl_xor(ad, j, 64);
        const int32_t ip = __shfl_xor(ai, j, 64);
        const bool mine_gt = ad > dp || (ad == dp && ai > ip);
        const bool lower = (lane & j) == 0;
        if (lower == mine_gt) {
            ad = dp;
            ai = ip;
        }
    }
}

// One wave per rescanned query of the round: merges its slices' lists and
// overwrites its k results (square roots; sentinels past the references).
__global__ __launch_bounds__(64) void knn_exact_merge_kernel(
    const float *ex_d, const int32_t *ex_i, const int32_t *ex_n, const int32_t *fail_list,
    const int32_t *fail_count, int64_t base, int slices, int k, int64_t nr, int32_t *out_i,
    float *out_d) {
    const int64_t slot = blockIdx.x;
    if (base + slot >= *fail_count) return;
    const int64_t q = fail_list[base + slot];
    const int lane = threadIdx.x;
    float ad = INFINITY;
    int32_t ai = INT32_MAX;
    int na = 0;
    for (int sl = 0; sl < slices; ++sl) {
        const int64_t cell = slot * slices + sl;
        const int nb = ex_n[cell];
        if (nb == 0) continue;
        const int r = kTop - 1 - lane;
        merge_reversed(ad, ai, r < nb ? ex_d[cell * kTop + r] : INFINITY,
                       r < nb ? ex_i[cell * kTop + r] : INT32_MAX, lane);
        na = min(kTop, na + nb);
    }
    if (lane < k) {
        out_i[q * k + lane] = lane < na ? ai : static_cast<int32_t>(nr);
        out_d[q * k + lane] = lane < na ? sqrtf(ad) : sqrtf(FLT_MAX);
    }
}

// 2 n u / (1 - n u) for f32 sums of n terms, doubled.
double sum_bound(int64_t n) {
    const double nu = static_cast<double>(n) * 0x1.0p-24;
    return 2.0 * nu / (1.0 - nu);
}

template <int KS>
void launch_scan(dim3 cells, hipStream_t s, const uint16_t *qb, int64_t nq, const uint16_t *rb,
                 const float *rnorm, int64_t nr, int64_t per_slice, int slices, float *cd,
                 int32_t *ci, int32_t *cn, uint32_t *tk, int exclude_self, int64_t qoff) {
    // cells.x query blocks by cells.y slices, as 8 * ceil(cells / 8) workgroups (scan_cell)
    const int64_t n_cells = static_cast<int64_t>(cells.x) * cells.y;
    const dim3 grid(static_cast<unsigned>(8 * ((n_cells + 7) / 8)));
    if (n_cells <= 256)
        hipLaunchKernelGGL((knn_scan_small_kernel<KS>), grid, dim3(256), 0, s, qb, nq, rb, rnorm, nr,
                           per_slice, slices, cd, ci, cn, tk, exclude_self, qoff);
    else if (exclude_self && nq <= 64)
        hipLaunchKernelGGL((knn_scan_kernel<KS, kScanRing, false, true, 1>), grid, dim3(256), 0, s, qb,
                           nq, rb, rnorm, nr, per_slice, slices, cd, ci, cn, tk, exclude_self, qoff,
                           int64_t{0}, static_cast<float *>(nullptr));
    else if (exclude_self)
        hipLaunchKernelGGL((knn_scan_kernel<KS, kScanRing, false, true, 2>), grid, dim3(256), 0, s, qb,
                           nq, rb, rnorm, nr, per_slice, slices, cd, ci, cn, tk, exclude_self, qoff,
                           int64_t{0}, static_cast<float *>(nullptr));
    else if (nq <= 64)
        hipLaunchKernelGGL((knn_scan_kernel<KS, kScanRing, false, false, 1>), grid, dim3(256), 0, s, qb,
                           nq, rb, rnorm, nr, per_slice, slices, cd, ci, cn, tk, exclude_self, qoff,
                           int64_t{0}, static_cast<float *>(nullptr));
    else
        hipLaunchKernelGGL((knn_scan_kernel<KS, kScanRing, false, false, 2>), grid, dim3(256), 0, s, qb,
                           nq, rb, rnorm, nr, per_slice, slices, cd, ci, cn, tk, exclude_self, qoff,
                           int64_t{0}, static_cast<float *>(nullptr));
}

template <int KS>
void launch_sample_dist(dim3 grid, hipStream_t s, const uint16_t *qb, int64_t nq, const uint16_t *rb,
                        const float *rnorm, int64_t nr, int m, float *dist, int exclude_self,
                        int64_t qoff) {
    // grid.x query blocks by grid.y sample slices of m / grid.y entries
    // each, through the scan's own pipeline (cells as in launch_scan)
    const int64_t n_cells = static_cast<int64_t>(grid.x) * grid.y;
    const dim3 wgs(static_cast<unsigned>(8 * ((n_cells + 7) / 8)));
    if (nq <= 64)
        hipLaunchKernelGGL((knn_scan_kernel<KS, kScanRing, true, true, 1>), wgs, dim3(256), 0, s, qb, nq,
                           rb, rnorm, nr, static_cast<int64_t>(m / grid.y), static_cast<int>(grid.y),
                           static_cast<float *>(nullptr), static_cast<int32_t *>(nullptr),
                           static_cast<int32_t *>(nullptr), static_cast<uint32_t *>(nullptr),
                           exclude_self, qoff, static_cast<int64_t>(m), dist);
    else
        hipLaunchKernelGGL((knn_scan_kernel<KS, kScanRing, true, true, 2>), wgs, dim3(256), 0, s, qb, nq,
                           rb, rnorm, nr, static_cast<int64_t>(m / grid.y), static_cast<int>(grid.y),
                           static_cast<float *>(nullptr), static_cast<int32_t *>(nullptr),
                           static_cast<int32_t *>(nullptr), static_cast<uint32_t *>(nullptr),
                           exclude_self, qoff, static_cast<int64_t>(m), dist);
}

int ksteps_for(int64_t d) {
    const int64_t need = (d + 31) / 32;
    const int choices[] = {1, 2, 4, 8, 16, 25, 32};
    for (int c : choices)
        if (c >= need) return c;
    return -1;
}

}  // namespace

// ------------------------------------------------------------- KnnIndex
KnnIndex::KnnIndex(const float *ref_host, const float *ref_device, int64_t n, int64_t d, int64_t max_q,
                   hipStream_t s)
    : n_(n), d_(d) {
    ks_ = ksteps_for(d);
    UMAPHIP_REQUIRE(ks_ > 0, UMAPHIP_INVALID_INPUT, "kNN supports at most 1024 features (got %lld)",
                    (long long)d);
    dp_ = static_cast<int64_t>(ks_) * 32;
    ref_.alloc(static_cast<size_t>(n * d));
    if (ref_device)
        HIP_CHECK(hipMemcpyAsync(ref_.ptr, ref_device, n * d * sizeof(float), hipMemcpyDeviceToDevice, s));
    else
        HIP_CHECK(hipMemcpyAsync(ref_.ptr, ref_host, n * d * sizeof(float), hipMemcpyHostToDevice, s));
    // the per-feature shift from the references' moments
    shift_.alloc(static_cast<size_t>(d));
    {
        const int64_t blocks = (n + kMomentRows - 1) / kMomentRows;
        DeviceArray<double> partial(static_cast<size_t>(2 * blocks * d));
        hipLaunchKernelGGL(column_moments_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256),
                           0, s, ref_.ptr, n, d, partial.ptr);
        hipLaunchKernelGGL(shift_kernel, dim3(static_cast<unsigned>((d + 255) / 256)), dim3(256),
                           0, s, partial.ptr, blocks, n, d, shift_.ptr);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(s));   // before the partials are freed
    }
    ref_bf_.alloc(static_cast<size_t>(n * dp_));
    rnorm_.alloc(static_cast<size_t>(n));
    eps_key_.alloc(1);
    {
        DeviceArray<float> rerr(static_cast<size_t>(n));
        hipLaunchKernelGGL(to_bf16_kernel, dim3(static_cast<unsigned>((n + 3) / 4)), dim3(64, 4), 0,
                           s, ref_.ptr, n, d, dp_, shift_.ptr, ref_bf_.ptr, rnorm_.ptr, rerr.ptr);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, eps_key_.ptr, 1, 0u);
        hipLaunchKernelGGL(rounding_ratio_kernel, dim3(static_cast<unsigned>((n + 255) / 256)),
                           dim3(256), 0, s, rnorm_.ptr, rerr.ptr, n, sum_bound(dp_), eps_key_.ptr);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(s));   // before the row errors are freed
    }
    ex_slices_ = static_cast<int>(std::min<int64_t>(64, std::max<int64_t>(1, (n + 8191) / 8192)));
    fail_count_.alloc(1);
    reserve(max_q);
}

bool KnnIndex::uses_sample(int64_t q) const {
    return n_ >= kSampleMinRefs && slices_for(q) >= kSampleMinSlices;
}

int KnnIndex::sample_size(int64_t q) const {
    // the larger sample (a bound about 4x tighter, so fewer survivors to
    // append, sort and merge) where its distances stay cheap next to the scan
    static_assert(kSample == 256 * kSelectKeys && kSampleLarge == 1024 * kSelectKeys,
                  "the select holds the sample in registers");
    return q <= kLargeSampleMaxQ && n_ >= 8 * static_cast<int64_t>(kSampleLarge) ? kSampleLarge
                                                                               : kSample;
}

int KnnIndex::slices_for(int64_t q) const {
    // Enough workgroups to fill the chip (256 CUs, two resident each), but
    // slices of at least kMinSliceRefs references. Rounded down: 513 cells
    // would run one cell in a second round and double the scan.
    const int64_t blocks = (q + kTQ - 1) / kTQ;
    int64_t s = 512 / blocks;
    s = std::min<int64_t>(s, std::max<int64_t>(1, n_ / kMinSliceRefs));
    s = std::max<int64_t>(1, std::min<int64_t>(s, kMaxSlices));
    return static_cast<int>(s);
}

void KnnIndex::reserve(int64_t q) {
    if (q <= reserved_q_) return;
    // (query, slice) buffers are laid out per batch; size them for the
    // largest rows x slices product of any batch up to q rows.
    int64_t cells = 0;
    for (int64_t qq = 1; qq <= q; ++qq) cells = std::max<int64_t>(cells, qq * slices_for(qq));
    const size_t entries = static_cast<size_t>(cells) * kCap;
    cand_d_.alloc(entries);
    cand_i_.alloc(entries);
    cand_n_.alloc(static_cast<size_t>(cells));
    top_d_.alloc(static_cast<size_t>(q) * kTop);
    top_i_.alloc(static_cast<size_t>(q) * kTop);
    top_n_.alloc(static_cast<size_t>(q));
    tau_key_.alloc(static_cast<size_t>(q));
    // sample distances of the largest batch that uses the sample bound
    int64_t sq = 0;
    for (int64_t qq = q; qq >= 1 && sq == 0; --qq)
        if (uses_sample(qq)) sq = qq;
    if (sq > 0)
        sample_d_.alloc(static_cast<size_t>(std::max<int64_t>(
            sq * kSample, std::min<int64_t>(sq, kLargeSampleMaxQ) * sample_size(1))));
    qbf_.alloc(static_cast<size_t>(q * dp_));
    qnorm_.alloc(static_cast<size_t>(q));
    qerr_.alloc(static_cast<size_t>(q));
    kth_.alloc(static_cast<size_t>(q));
    fail_list_.alloc(static_cast<size_t>(q));
    const int64_t round = std::min<int64_t>(q, kExactRound);
    ex_d_.alloc(static_cast<size_t>(round * ex_slices_ * kTop));
    ex_i_.alloc(static_cast<size_t>(round * ex_slices_ * kTop));
    ex_n_.alloc(static_cast<size_t>(round * ex_slices_));
    reserved_q_ = q;
}

void KnnIndex::query(const float *q_dev, int64_t nq, int k, bool exclude_self, int64_t query_offset,
                     int32_t *idx_out, float *dist_out, hipStream_t s, const float *q_f32_for_rerank,
                     int32_t *cand_ids, int32_t *cand_counts) {
    UMAPHIP_REQUIRE(k >= 1 && k <= kTop, UMAPHIP_INVALID_INPUT, "k must be in 1..=%d", kTop);
    if (nq == 0) return;
    UMAPHIP_REQUIRE(nq <= reserved_q_, UMAPHIP_CAPACITY,
                    "query batch %lld exceeds the reserved %lld rows", (long long)nq,
                    (long long)reserved_q_);
    hipLaunchKernelGGL(to_bf16_kernel, dim3(static_cast<unsigned>((nq + 3) / 4)), dim3(64, 4), 0, s,
                       q_dev, nq, d_, dp_, shift_.ptr, qbf_.ptr, qnorm_.ptr, qerr_.ptr);
    HIP_CHECK(hipGetLastError());
    const int slices = slices_for(nq);
    const int64_t per_slice = ((n_ + slices - 1) / slices + kTR - 1) / kTR * kTR;
    const dim3 grid(static_cast<unsigned>((nq + kTQ - 1) / kTQ), static_cast<unsigned>(slices));
    const int ex = exclude_self ? 1 : 0;
    if (uses_sample(nq)) {
        // every query's global threshold starts at its sample bound; about
        // 512 workgroups, each loading its queries' fragments once for
        // m / (sample slices) references
        const int m = sample_size(nq);
        const int64_t qblocks = (nq + kTQ - 1) / kTQ;
        int sample_slices = m / kSampleRefs;
        while (sample_slices > 1 && qblocks * sample_slices > 512) sample_slices /= 2;
        const dim3 sgrid(static_cast<unsigned>(qblocks), static_cast<unsigned>(sample_slices));
        switch (ks_) {
            case 1: launch_sample_dist<1>(sgrid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, m, sample_d_.ptr, ex, query_offset); break;
            case 2: launch_sample_dist<2>(sgrid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, m, sample_d_.ptr, ex, query_offset); break;
            case 4: launch_sample_dist<4>(sgrid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, m, sample_d_.ptr, ex, query_offset); break;
            case 8: launch_sample_dist<8>(sgrid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, m, sample_d_.ptr, ex, query_offset); break;
            case 16: launch_sample_dist<16>(sgrid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, m, sample_d_.ptr, ex, query_offset); break;
            case 25: launch_sample_dist<25>(sgrid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, m, sample_d_.ptr, ex, query_offset); break;
            default: launch_sample_dist<32>(sgrid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, m, sample_d_.ptr, ex, query_offset); break;
        }
        HIP_CHECK(hipGetLastError());
        if (m == kSampleLarge)
            hipLaunchKernelGGL(knn_sample_select_kernel<16>, dim3(static_cast<unsigned>(nq)), dim3(1024),
                               0, s, sample_d_.ptr, m, tau_key_.ptr);
        else
            hipLaunchKernelGGL(knn_sample_select_kernel<4>, dim3(static_cast<unsigned>(nq)), dim3(256), 0,
                               s, sample_d_.ptr, m, tau_key_.ptr);
    } else {
        hipLaunchKernelGGL(fill_u32_kernel, dim3(static_cast<unsigned>((nq + 255) / 256)), dim3(256), 0,
                           s, tau_key_.ptr, nq, 0xFF800000u);   // key of +inf
    }
    HIP_CHECK(hipGetLastError());
    switch (ks_) {
        case 1: launch_scan<1>(grid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, per_slice, slices, cand_d_.ptr, cand_i_.ptr, cand_n_.ptr, tau_key_.ptr, ex, query_offset); break;
        case 2: launch_scan<2>(grid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, per_slice, slices, cand_d_.ptr, cand_i_.ptr, cand_n_.ptr, tau_key_.ptr, ex, query_offset); break;
        case 4: launch_scan<4>(grid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, per_slice, slices, cand_d_.ptr, cand_i_.ptr, cand_n_.ptr, tau_key_.ptr, ex, query_offset); break;
        case 8: launch_scan<8>(grid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, per_slice, slices, cand_d_.ptr, cand_i_.ptr, cand_n_.ptr, tau_key_.ptr, ex, query_offset); break;
        case 16: launch_scan<16>(grid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, per_slice, slices, cand_d_.ptr, cand_i_.ptr, cand_n_.ptr, tau_key_.ptr, ex, query_offset); break;
        case 25: launch_scan<25>(grid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, per_slice, slices, cand_d_.ptr, cand_i_.ptr, cand_n_.ptr, tau_key_.ptr, ex, query_offset); break;
        default: launch_scan<32>(grid, s, qbf_.ptr, nq, ref_bf_.ptr, rnorm_.ptr, n_, per_slice, slices, cand_d_.ptr, cand_i_.ptr, cand_n_.ptr, tau_key_.ptr, ex, query_offset); break;
    }
    HIP_CHECK(hipGetLastError());
    if (nq <= kDeepMergeMaxQ)
        hipLaunchKernelGGL(knn_merge_kernel<16>, dim3(static_cast<unsigned>(nq)), dim3(1024), 0, s,
                           cand_d_.ptr, cand_i_.ptr, cand_n_.ptr, tau_key_.ptr, nq, slices,
                           top_d_.ptr, top_i_.ptr, top_n_.ptr, fail_count_.ptr);
    else if (nq <= kWideMergeMaxQ)
        hipLaunchKernelGGL(knn_merge_kernel<4>, dim3(static_cast<unsigned>(nq)), dim3(256), 0, s,
                           cand_d_.ptr, cand_i_.ptr, cand_n_.ptr, tau_key_.ptr, nq, slices,
                           top_d_.ptr, top_i_.ptr, top_n_.ptr, fail_count_.ptr);
    else
        hipLaunchKernelGGL(knn_merge_kernel<1>, dim3(static_cast<unsigned>(nq)), dim3(64), 0, s,
                           cand_d_.ptr, cand_i_.ptr, cand_n_.ptr, tau_key_.ptr, nq, slices,
                           top_d_.ptr, top_i_.ptr, top_n_.ptr, fail_count_.ptr);
    HIP_CHECK(hipGetLastError());
    if (cand_ids) {
        // prefilter lists only (best first), no rerank
        HIP_CHECK(hipMemcpyAsync(cand_ids, top_i_.ptr, nq * kTop * sizeof(int32_t),
                                 hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync(cand_counts, top_n_.ptr, nq * sizeof(int32_t),
                                 hipMemcpyDeviceToDevice, s));
        return;
    }
    // one four-wave workgroup per query: it also has more loads in flight
    // per CU at large batches (87 ns per query with one wave at 16,384
    // rows); small batches load each query's rows in one round trip
    // with each query's certificate (the merge cleared the open count)
    const float *q_exact = q_f32_for_rerank ? q_f32_for_rerank : q_dev;
    const CertifyArgs ca{top_d_.ptr,      qnorm_.ptr,         qerr_.ptr,      eps_key_.ptr,
                         sum_bound(dp_), force_rescan_ ? 1 : 0, fail_list_.ptr, fail_count_.ptr};
    if (nq <= kRerankWideMaxQ && d_ <= 64 * kRerankSlots)
        hipLaunchKernelGGL(knn_rerank_wide_kernel, dim3(static_cast<unsigned>(nq)), dim3(1024), 0, s,
                           q_exact, nq, ref_.ptr, n_, d_, top_i_.ptr, top_n_.ptr, k, idx_out,
                           dist_out, kth_.ptr, ca);
    else
        hipLaunchKernelGGL(knn_rerank_kernel, dim3(static_cast<unsigned>(nq)), dim3(256), 0, s,
                           q_exact, nq, ref_.ptr, n_, d_, top_i_.ptr, top_n_.ptr, k, idx_out,
                           dist_out, kth_.ptr, ca);
    HIP_CHECK(hipGetLastError());
    // the exact rescan of the queries the certificate leaves open
    const int64_t round = std::min<int64_t>(nq, kExactRound);
    int64_t open = nq;
    if (nq > kWideMergeMaxQ) {
        // large batches (not latency bound) launch only the rounds needed
        int32_t host = 0;
        HIP_CHECK(hipMemcpyAsync(&host, fail_count_.ptr, sizeof(host), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        open = host;
    }
    const int64_t ex_per_slice = ((n_ + ex_slices_ - 1) / ex_slices_ + 63) / 64 * 64;
    for (int64_t base = 0; base < open; base += round) {
        const dim3 sgrid(static_cast<unsigned>((round + kExactQB - 1) / kExactQB),
                         static_cast<unsigned>(ex_slices_));
        hipLaunchKernelGGL(knn_exact_scan_kernel, sgrid, dim3(64), 0, s, q_exact, ref_.ptr, n_, d_,
                           fail_list_.ptr, fail_count_.ptr, base, ex_per_slice, ex_slices_, kth_.ptr,
                           exclude_self ? 1 : 0, query_offset, ex_d_.ptr, ex_i_.ptr, ex_n_.ptr);
        hipLaunchKernelGGL(knn_exact_merge_kernel, dim3(static_cast<unsigned>(round)), dim3(64), 0,
                           s, ex_d_.ptr, ex_i_.ptr, ex_n_.ptr, fail_list_.ptr, fail_count_.ptr, base,
                           ex_slices_, k, n_, idx_out, dist_out);
        HIP_CHECK(hipGetLastError());
    }
}

int64_t KnnIndex::rescans(hipStream_t s) const {
    int32_t host = 0;
    HIP_CHECK(hipMemcpyAsync(&host, fail_count_.ptr, sizeof(host), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return host;
}

float KnnIndex::rounding_ratio(hipStream_t s) const {
    uint32_t host = 0;
    HIP_CHECK(hipMemcpyAsync(&host, eps_key_.ptr, sizeof(host), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    float v;
    std::memcpy(&v, &host, sizeof(v));
    return v;
}

}  // namespace umaphip

using namespace umaphip;

struct umaphip_knn_index {
    std::unique_ptr<KnnIndex> impl;
};

extern "C" {

umaphip_status umaphip_knn_index_create(const float *reference, int32_t reference_on_device,
                                        int64_t n, int64_t d, int64_t max_queries,
                                        umaphip_knn_index **out) {
    return guarded([&] {
        UMAPHIP_REQUIRE(out && reference && n > 0 && d > 0 && max_queries > 0 &&
                            n <= INT32_MAX - 1,
                        UMAPHIP_INVALID_INPUT, "invalid kNN index arguments");
        std::unique_ptr<umaphip_knn_index> idx(new umaphip_knn_index());
        idx->impl.reset(new KnnIndex(reference_on_device ? nullptr : reference,
                                     reference_on_device ? reference : nullptr, n, d, max_queries,
                                     nullptr));
        HIP_CHECK(hipDeviceSynchronize());
        *out = idx.release();
    });
}

umaphip_status umaphip_knn_query(umaphip_knn_index *index, const float *queries, int64_t q,
                                 int32_t k, int32_t exclude_self, int64_t query_offset,
                                 int32_t *indices, float *distances, void *stream) {
    return guarded([&] {
        UMAPHIP_REQUIRE(index && index->impl, UMAPHIP_INVALID_INPUT, "kNN index is null");
        UMAPHIP_REQUIRE(q >= 0 && queries && indices && distances, UMAPHIP_INVALID_INPUT,
                        "invalid kNN query arguments");
        index->impl->query(queries, q, k, exclude_self != 0, query_offset, indices, distances,
                           as_stream(stream), nullptr);
    });
}

umaphip_status umaphip_knn_candidates(umaphip_knn_index *index, const float *queries, int64_t q,
                                      int32_t *candidates, int32_t *counts, void *stream) {
    return guarded([&] {
        UMAPHIP_REQUIRE(index && index->impl, UMAPHIP_INVALID_INPUT, "kNN index is null");
        UMAPHIP_REQUIRE(q >= 0 && queries && candidates && counts, UMAPHIP_INVALID_INPUT,
                        "invalid kNN candidate arguments");
        index->impl->query(queries, q, 1, false, 0, nullptr, nullptr, as_stream(stream), nullptr,
                           candidates, counts);
    });
}

umaphip_status umaphip_knn_set_rescan(umaphip_knn_index *index, int32_t every_query) {
    return guarded([&] {
        UMAPHIP_REQUIRE(index && index->impl, UMAPHIP_INVALID_INPUT, "kNN index is null");
        index->impl->set_force_rescan(every_query != 0);
    });
}

umaphip_status umaphip_knn_certificate_stats(umaphip_knn_index *index, void *stream,
                                             int64_t *rescanned, float *rounding_ratio) {
    return guarded([&] {
        UMAPHIP_REQUIRE(index && index->impl, UMAPHIP_INVALID_INPUT, "kNN index is null");
        if (rescanned) *rescanned = index->impl->rescans(as_stream(stream));
        if (rounding_ratio) *rounding_ratio = index->impl->rounding_ratio(as_stream(stream));
    });
}

void umaphip_knn_index_destroy(umaphip_knn_index *index) { delete index; }

}  // extern "C"
