// Host-side graph and initialisation steps that feed the layout:
// the fuzzy simplicial set (smooth kNN + symmetrisation), positive-edge
// pruning and compaction into the layout CSR with f64 clocks, the curve fit
// for (a, b), random initialisation with isotropic rescaling, and the
// transform's weighted initialisation. All are multi-threaded host C++;
// the epoch loop they feed runs on the device.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "common.hpp"
#include "host_parallel.hpp"
#include "stdrng.hpp"

namespace umaphip {
namespace {

// smooth_row (crates/umap/src/kernels/standard_umap/graph.rs:117-242).
void smooth_row(const float *dist, int64_t k, float global_mean, uint32_t floor_conn,
                float fraction, uint32_t min_positive, bool interpolate, float target, float &rho_out,
                float &sigma_out) {
    int64_t positives = 0;
    float pmax = 0.0f;
    for (int64_t s = 0; s < k; ++s)
        if (dist[s] > 0.0f) {
            ++positives;
            pmax = std::max(pmax, dist[s]);
        }
    float rho = 0.0f;
    if (positives > 0) {
        if (floor_conn == 0) {
            float first = pmax;
            for (int64_t s = 0; s < k; ++s)
                if (dist[s] > 0.0f && dist[s] < first) first = dist[s];
            rho = first * fraction;
        } else {
            const int64_t lo_rank = static_cast<int64_t>(floor_conn) - 1, hi_rank = floor_conn;
            float before = pmax, after = pmax;
            for (int64_t c = 0; c < k; ++c) {
                const float cand = dist[c];
                if (!(cand > 0.0f)) continue;
                int64_t less = 0, equal = 0;
                for (int64_t o = 0; o < k; ++o) {
                    if (dist[o] > 0.0f && dist[o] < cand) ++less;
                    else if (dist[o] > 0.0f && dist[o] == cand) ++equal;
                }
                if (less <= lo_rank && lo_rank < less + equal) before = cand;
                if (less <= hi_rank && hi_rank < less + equal) after = cand;
            }
            rho = interpolate ? before + (after - before) * fraction : before;
        }
        if (positives < static_cast<int64_t>(min_positive)) rho = pmax;
    }
    float row_sum = 0.0f;
    for (int64_t s = 0; s < k; ++s) row_sum += dist[s];
    const float floor_mean = rho > 0.0f ? row_sum / static_cast<float>(static_cast<uint32_t>(k)) : global_mean;
    float lo = 0.0f, hi = 3.4028235e38f, mid = 1.0f;
    for (int it = 0; it < 64; ++it) {
        float sum = 0.0f;
        for (int64_t s = 1; s < k; ++s) {
            const float delta = dist[s] - rho;
            const float safe = mid == 0.0f ? 1.0f : mid;
            const float clamped = delta < 0.0f ? 0.0f : delta;
            float value = std::exp(-clamped / safe);
            if (delta <= 0.0f) value = 1.0f;
            sum += value;
        }
        if (std::fabs(sum - target) < 1e-5f) break;
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
    const float floor = floor_mean * 1e-3f;
    if (mid < floor) mid = floor;
    rho_out = rho;
    sigma_out = mid;
}

struct Emit {
    int32_t h1, t1, h2, t2;
    float w;
    int dirs;
};

}  // namespace
}  // namespace umaphip

using namespace umaphip;

extern "C" {

// fuzzy_graph (graph.rs:21-43, 252-564): memberships, symmetrisation and the
// per-head tail-sorted COO (capacity 2*n*k, zero tail, -2/-2/-1 marker on
// invalid neighbour ids). The global mean that floors sigma is summed in
// float64 (the reference leaves its f32 summation order to the backend).
umaphip_status umaphip_fuzzy_graph(const int32_t *indices, const float *distances, int64_t n,
                                   int64_t k, float local_connectivity, float bandwidth, float mix,
                                   int32_t bipartite, int64_t n_reference, int32_t *heads,
                                   int32_t *tails, float *weights, float *rho, float *sigma,
                                   float *directed, int64_t *used_out) {
    return guarded([&] {
        UMAPHIP_REQUIRE(indices && distances && heads && tails && weights && rho && sigma &&
                            directed && used_out && n > 0 && k > 0,
                        UMAPHIP_INVALID_INPUT, "invalid fuzzy graph arguments");
        UMAPHIP_REQUIRE(std::isfinite(local_connectivity) && local_connectivity >= 0.0f &&
                            std::isfinite(bandwidth) && bandwidth > 0.0f && std::isfinite(mix) &&
                            mix >= 0.0f && mix <= 1.0f,
                        UMAPHIP_INVALID_CONFIG, "invalid fuzzy graph settings");
        UMAPHIP_REQUIRE(bipartite || n == n_reference, UMAPHIP_INVALID_INPUT,
                        "fit graph reference count must equal query count");
        UMAPHIP_REQUIRE(2 * n * k <= INT32_MAX, UMAPHIP_OVERFLOW, "fuzzy graph COO indices");
        const uint64_t floor_c = static_cast<uint64_t>(local_connectivity);
        const float fraction = local_connectivity - static_cast<float>(floor_c);
        const uint64_t min_pos = floor_c + (local_connectivity > static_cast<float>(floor_c) ? 1 : 0);
        const uint64_t above = static_cast<uint64_t>(k) + 1;
        const uint32_t fc = static_cast<uint32_t>(std::min(floor_c, above));
        const uint32_t mp = static_cast<uint32_t>(std::min(min_pos, above));
        const bool interpolate = fraction > 1e-5f;
        const float target = std::log2(static_cast<float>(k)) * bandwidth;
        double total = 0.0;
        for (int64_t i = 0; i < n * k; ++i) total += distances[i];
        const float global_mean = static_cast<float>(total / static_cast<double>(n * k));
        std::atomic<int> invalid{0};
        parallel_for(n, [&](int64_t row) {
            float r, s;
            smooth_row(distances + row * k, k, global_mean, fc, fraction, mp, interpolate, target, r, s);
            rho[row] = r;
            sigma[row] = s;
            for (int64_t slot = 0; slot < k; ++slot) {
                const int32_t index = indices[row * k + slot];
                const float dd = distances[row * k + slot];
                const float delta = dd - r;
                const float safe = s == 0.0f ? 1.0f : s;
                const float clamped = delta < 0.0f ? 0.0f : delta;
                float value = std::exp(-clamped / safe);
                if (dd <= r || s == 0.0f) value = 1.0f;
                if (index < 0 || index >= n_reference || (!bipartite && index == row)) value = 0.0f;
                directed[row * k + slot] = value;
                if (index < -1 || index >= n_reference) invalid.store(1);
            }
        }, 256);
        // evaluate_edge (graph.rs:252-337)
        std::vector<Emit> ev(static_cast<size_t>(n * k));
        const int64_t nodes = bipartite ? n + n_reference : n;
        std::vector<std::atomic<int64_t>> counts(static_cast<size_t>(nodes) + 1);
        for (auto &c : counts) c.store(0, std::memory_order_relaxed);
        parallel_for(n, [&](int64_t row) {
            for (int64_t slot = 0; slot < k; ++slot) {
                Emit &E = ev[row * k + slot];
                E.dirs = 0;
                const int32_t tail = indices[row * k + slot];
                if (!(tail >= 0 && tail < n_reference && (bipartite || tail != row))) continue;
                bool first = true;
                float forward = 0.0f;
                for (int64_t o = 0; o < k; ++o)
                    if (indices[row * k + o] == tail) {
                        if (o < slot) first = false;
                        forward += directed[row * k + o];
                    }
                if (!first) continue;
                float reverse = 0.0f;
                bool has_reverse = false;
                if (!bipartite)
                    for (int64_t rs = 0; rs < k; ++rs)
                        if (indices[static_cast<int64_t>(tail) * k + rs] == static_cast<int32_t>(row)) {
                            reverse += directed[static_cast<int64_t>(tail) * k + rs];
                            has_reverse = true;
                        }
                if (!(!has_reverse || row <= tail)) continue;
                float fuzzy;
                if (bipartite) {
                    fuzzy = forward;
                } else {
                    const float inter = forward * reverse;
                    const float uni = (forward + reverse) - inter;
                    fuzzy = mix * uni + (1.0f - mix) * inter;
                }
                if (!(fuzzy > 0.0f)) continue;
                const int32_t r32 = static_cast<int32_t>(row);
                const bool fwd = bipartite || r32 < tail;
                E.h1 = fwd ? r32 : tail;
                E.t1 = fwd ? tail : r32;
                E.h2 = E.t1;
                E.t2 = E.h1;
                E.w = fuzzy;
                E.dirs = bipartite ? 1 : 2;
                counts[E.h1 + 1].fetch_add(1, std::memory_order_relaxed);
                if (E.dirs > 1) counts[E.h2 + 1].fetch_add(1, std::memory_order_relaxed);
            }
        }, 256);
        std::vector<int64_t> offs(static_cast<size_t>(nodes) + 1, 0);
        for (int64_t i = 0; i < nodes; ++i) offs[i + 1] = offs[i] + counts[i + 1].load();
        std::vector<std::atomic<int64_t>> cursor(static_cast<size_t>(nodes));
        for (int64_t i = 0; i < nodes; ++i) cursor[i].store(offs[i], std::memory_order_relaxed);
        const int64_t capacity = 2 * n * k;
        std::fill(heads, heads + capacity, 0);
        std::fill(tails, tails + capacity, 0);
        std::fill(weights, weights + capacity, 0.0f);
        parallel_for(n * k, [&](int64_t e) {
            const Emit &E = ev[e];
            if (!E.dirs) return;
            int64_t p = cursor[E.h1].fetch_add(1, std::memory_order_relaxed);
            heads[p] = E.h1; tails[p] = E.t1; weights[p] = E.w;
            if (E.dirs > 1) {
                p = cursor[E.h2].fetch_add(1, std::memory_order_relaxed);
                heads[p] = E.h2; tails[p] = E.t2; weights[p] = E.w;
            }
        }, 4096);
        // per-head segments sorted by tail (tails are unique per head)
        parallel_for(nodes, [&](int64_t hd) {
            const int64_t b = offs[hd], e = offs[hd + 1];
            if (e - b < 2) return;
            std::vector<std::pair<int32_t, float>> seg;
            seg.reserve(static_cast<size_t>(e - b));
            for (int64_t p = b; p < e; ++p) seg.emplace_back(tails[p], weights[p]);
            std::sort(seg.begin(), seg.end(),
                      [](const auto &x, const auto &y) { return x.first < y.first; });
            for (int64_t p = b; p < e; ++p) {
                tails[p] = seg[p - b].first;
                weights[p] = seg[p - b].second;
            }
        }, 1024);
        if (invalid.load() && capacity > 0) {
            heads[0] = -2;
            tails[0] = -2;
            weights[0] = -1.0f;
        }
        *used_out = offs[nodes];
        UMAPHIP_REQUIRE(!invalid.load(), UMAPHIP_NUMERICAL,
                        "fuzzy graph contains missing or invalid neighbor indices");
    });
}

// prepare_schedule pruning (crates/umap/src/standard_umap/estimator.rs:2753-
// 2798) followed by compact_graph (crates/umap-core/src/modern_layout.rs:
// 188-208, 303-430). Zero weights are dropped; duplicates are summed in
// input order; rows are sorted by column. prune_divisor <= 0 skips pruning.
umaphip_status umaphip_compact_graph(int64_t rows, int64_t columns, int64_t n_edges,
                                     const int32_t *heads, const int32_t *tails,
                                     const float *weights_in, double prune_divisor, int32_t epochs,
                                     float rate, int32_t transform, uint32_t *offsets,
                                     uint32_t *indices, float *weights, double *periods,
                                     double *negative_periods, double *next_positive,
                                     double *next_negative, int64_t *n_compact) {
    return guarded([&] {
        UMAPHIP_REQUIRE(rows > 0 && columns > 0 && n_edges >= 0 && offsets && n_compact,
                        UMAPHIP_INVALID_INPUT, "invalid compaction arguments");
        UMAPHIP_REQUIRE(rows <= UINT32_MAX && columns <= UINT32_MAX && n_edges <= UINT32_MAX,
                        UMAPHIP_OVERFLOW, "CSR sizes");
        UMAPHIP_REQUIRE(std::isfinite(rate) && rate >= 0.0f && epochs > 0, UMAPHIP_INVALID_CONFIG,
                        "epochs must be positive and negative sample rate finite and nonnegative");
        float maximum = 0.0f;
        for (int64_t e = 0; e < n_edges; ++e) {
            const float w = weights_in[e];
            UMAPHIP_REQUIRE(std::isfinite(w) && w >= 0.0f, UMAPHIP_INVALID_INPUT,
                            "edge weights must be finite and nonnegative");
            maximum = std::max(maximum, w);
        }
        const float threshold =
            prune_divisor > 0.0 ? maximum / static_cast<float>(prune_divisor) : 0.0f;
        // counting sort by row, stable in input order
        std::vector<int64_t> count(static_cast<size_t>(rows) + 1, 0);
        for (int64_t e = 0; e < n_edges; ++e) {
            const float w = weights_in[e];
            if (w == 0.0f || w < threshold) continue;
            UMAPHIP_REQUIRE(heads[e] >= 0 && heads[e] < rows && tails[e] >= 0 && tails[e] < columns,
                            UMAPHIP_INVALID_INPUT, "edge index outside CSR shape");
            ++count[heads[e] + 1];
        }
        for (int64_t r = 0; r < rows; ++r) count[r + 1] += count[r];
        const int64_t kept = count[rows];
        std::vector<int64_t> order(static_cast<size_t>(kept));
        {
            std::vector<int64_t> cur(count.begin(), count.end() - 1);
            for (int64_t e = 0; e < n_edges; ++e) {
                const float w = weights_in[e];
                if (w == 0.0f || w < threshold) continue;
                order[cur[heads[e]]++] = e;
            }
        }
        // per row: stable sort by column, merge duplicates
        std::vector<int64_t> row_len(static_cast<size_t>(rows), 0);
        std::vector<uint32_t> cols(static_cast<size_t>(kept));
        std::vector<float> wts(static_cast<size_t>(kept));
        std::atomic<int> overflow{0};
        parallel_for(rows, [&](int64_t r) {
            const int64_t b = count[r], e = count[r + 1];
            std::stable_sort(order.begin() + b, order.begin() + e,
                             [&](int64_t x, int64_t y) { return tails[x] < tails[y]; });
            int64_t m = b;
            for (int64_t p = b; p < e; ++p) {
                const uint32_t c = static_cast<uint32_t>(tails[order[p]]);
                const float w = weights_in[order[p]];
                if (m > b && cols[m - 1] == c) {
                    wts[m - 1] = wts[m - 1] + w;
                    if (!std::isfinite(wts[m - 1])) overflow.store(1);
                    continue;
                }
                cols[m] = c;
                wts[m] = w;
                ++m;
            }
            row_len[r] = m - b;
        }, 512);
        UMAPHIP_REQUIRE(!overflow.load(), UMAPHIP_NUMERICAL, "duplicate edge weight overflow");
        offsets[0] = 0;
        for (int64_t r = 0; r < rows; ++r) offsets[r + 1] = offsets[r] + static_cast<uint32_t>(row_len[r]);
        const int64_t E = offsets[rows];
        UMAPHIP_REQUIRE(E > 0, UMAPHIP_INVALID_INPUT,
                        transform ? "transform fuzzy graph is empty after positive-edge pruning"
                                  : "fuzzy graph is empty after positive-edge pruning");
        float wmax = 0.0f;
        for (int64_t r = 0; r < rows; ++r)
            for (int64_t p = count[r]; p < count[r] + row_len[r]; ++p) wmax = std::max(wmax, wts[p]);
        std::atomic<int> bad{0};
        parallel_for(rows, [&](int64_t r) {
            const int64_t src = count[r];
            const int64_t dst = offsets[r];
            for (int64_t j = 0; j < row_len[r]; ++j) {
                const float w = wts[src + j];
                const float samples = static_cast<float>(epochs) * (w / wmax);
                const double period = static_cast<double>(epochs) / static_cast<double>(samples);
                const double nperiod =
                    rate == 0.0f ? INFINITY : (period / static_cast<double>(rate)) * 1.5;
                if (!std::isfinite(period) || period <= 0.0 ||
                    (rate > 0.0f && (!std::isfinite(nperiod) || nperiod <= 0.0)))
                    bad.store(1);
                indices[dst + j] = cols[src + j];
                weights[dst + j] = w;
                periods[dst + j] = period;
                negative_periods[dst + j] = nperiod;
                next_positive[dst + j] = period;
                next_negative[dst + j] = nperiod;
            }
        }, 4096);
        UMAPHIP_REQUIRE(!bad.load(), UMAPHIP_NUMERICAL, "invalid layout period");
        *n_compact = E;
    });
}

// fit_curve (crates/umap-core/src/config.rs:164-330): damped Gauss-Newton
// (Levenberg-Marquardt) on (log a, log b), 300 samples of [0, 3*spread].
umaphip_status umaphip_fit_curve(float min_dist, float spread, float *a_out, float *b_out) {
    return guarded([&] {
        UMAPHIP_REQUIRE(std::isfinite(min_dist) && min_dist >= 0.0f, UMAPHIP_INVALID_CONFIG,
                        "min_dist must be finite and nonnegative");
        UMAPHIP_REQUIRE(std::isfinite(spread) && spread > 0.0f, UMAPHIP_INVALID_CONFIG,
                        "spread must be finite and greater than zero");
        double xs[300], ts[300];
        for (int i = 0; i < 300; ++i) {
            xs[i] = 3.0 * static_cast<double>(spread) * i / 299.0;
            ts[i] = xs[i] < min_dist ? 1.0 : std::exp(-(xs[i] - min_dist) / static_cast<double>(spread));
        }
        auto curve = [](double x, double a, double b) {
            return x == 0.0 ? 1.0 : 1.0 / (1.0 + a * std::exp(2.0 * b * std::log(x)));
        };
        auto objective = [&](double la, double lb) {
            const double a = std::exp(la), b = std::exp(lb);
            double s = 0.0;
            for (int i = 0; i < 300; ++i) {
                const double r = curve(xs[i], a, b) - ts[i];
                s += r * r;
            }
            return s;
        };
        double la = 0.0, lb = 0.0, damping = 1e-3, obj = objective(la, lb);
        bool converged = false;
        for (int it = 0; it < 500; ++it) {
            const double a = std::exp(la), b = std::exp(lb);
            double haa = 0, hab = 0, hbb = 0, ga = 0, gb = 0;
            for (int i = 0; i < 300; ++i) {
                const double r = curve(xs[i], a, b) - ts[i];
                double ja = 0, jb = 0;
                if (xs[i] != 0.0) {
                    const double scaled = a * std::exp(2.0 * b * std::log(xs[i]));
                    const double dv = -scaled / ((1.0 + scaled) * (1.0 + scaled));
                    ja = dv;
                    jb = dv * 2.0 * b * std::log(xs[i]);
                }
                haa += ja * ja; hab += ja * jb; hbb += jb * jb; ga += ja * r; gb += jb * r;
            }
            if (std::max(std::fabs(ga), std::fabs(gb)) <= 1e-8) { converged = true; break; }
            const double daa = haa + damping * std::max(haa, 1.0);
            const double dbb = hbb + damping * std::max(hbb, 1.0);
            const double det = daa * dbb - hab * hab;
            if (!std::isfinite(det) || det <= 0.0) {
                damping *= 10.0;
                if (!std::isfinite(damping) || damping > 1e30) break;
                continue;
            }
            const double da = (dbb * ga - hab * gb) / det, db = (daa * gb - hab * ga) / det;
            UMAPHIP_REQUIRE(std::isfinite(da) && std::isfinite(db), UMAPHIP_NUMERICAL,
                            "curve fitting produced a non-finite parameter step");
            const double ca = la - da, cb = lb - db, cobj = objective(ca, cb);
            if (cobj < obj) {
                la = ca; lb = cb; obj = cobj;
                damping = std::max(damping / 3.0, 1e-15);
                if (std::max(std::fabs(da), std::fabs(db)) <= 1e-11) { converged = true; break; }
            } else {
                damping *= 10.0;
                if (!std::isfinite(damping) || damping > 1e30) break;
            }
        }
        const double a = std::exp(la), b = std::exp(lb);
        UMAPHIP_REQUIRE(converged && std::isfinite(a) && std::isfinite(b) && a > 0 && b > 0,
                        UMAPHIP_NUMERICAL, "curve fitting failed to converge");
        *a_out = static_cast<float>(a);
        *b_out = static_cast<float>(b);
    });
}

namespace {

// Order key of f32::total_cmp for non-NaN values (-0 before +0).
inline int32_t total_key(float x) {
    int32_t b;
    std::memcpy(&b, &x, 4);
    return b ^ static_cast<int32_t>(static_cast<uint32_t>(b >> 31) >> 1);
}

// init_rescale::rescale (crates/umap-core/src/init_rescale.rs:77-203):
// centre each axis (sequential f32 sum), then scale by the isotropic
// geometric mean of the 5-95% spans (mode 0), per axis (mode 1), or by the
// weight-median edge length of the graph (mode 2; edges in CSR order).
void rescale(float *v, int64_t n, int64_t d, const int64_t *offsets, const int32_t *indices,
             const float *weights, int32_t mode, float repulsion, float scale_factor) {
    auto finite = [](float x) { return std::isfinite(x); };
    for (int64_t c = 0; c < d; ++c) {
        float sum = 0.0f;
        for (int64_t r = 0; r < n; ++r) {
            UMAPHIP_REQUIRE(finite(v[r * d + c]), UMAPHIP_NUMERICAL, "non-finite initialization");
            sum += v[r * d + c];
            UMAPHIP_REQUIRE(finite(sum), UMAPHIP_NUMERICAL, "non-finite initialization");
        }
        const float mean = sum / static_cast<float>(n);
        for (int64_t r = 0; r < n; ++r) {
            v[r * d + c] -= mean;
            UMAPHIP_REQUIRE(finite(v[r * d + c]), UMAPHIP_NUMERICAL, "non-finite initialization");
        }
    }
    auto apply = [&](double multiplier) {
        const float m = static_cast<float>(multiplier);
        UMAPHIP_REQUIRE(finite(m), UMAPHIP_NUMERICAL, "non-finite initialization");
        for (int64_t i = 0; i < n * d; ++i) {
            v[i] *= m;
            UMAPHIP_REQUIRE(finite(v[i]), UMAPHIP_NUMERICAL, "non-finite initialization");
        }
    };
    if (mode == 2) {
        const int64_t E = offsets ? offsets[n] : 0;
        if (E == 0) return;
        struct Rec {
            float dist;
            float w;
            int64_t ordinal;
        };
        std::vector<Rec> recs(static_cast<size_t>(E));
        std::atomic<int> bad{0};
        parallel_for(n, [&](int64_t h) {
            for (int64_t e = offsets[h]; e < offsets[h + 1]; ++e) {
                const int64_t t = indices[e];
                float sq = 0.0f;
                for (int64_t c = 0; c < d; ++c) {
                    const float diff = v[h * d + c] - v[t * d + c];
                    sq += diff * diff;
                }
                const float dist = std::sqrt(sq);
                if (!finite(dist)) bad.store(1);
                recs[e] = {dist, weights[e], e};
            }
        }, 2048);
        UMAPHIP_REQUIRE(!bad.load(), UMAPHIP_NUMERICAL, "non-finite initialization");
        double total = 0.0;
        for (int64_t e = 0; e < E; ++e) {
            total += static_cast<double>(weights[e]);
            UMAPHIP_REQUIRE(std::isfinite(total), UMAPHIP_NUMERICAL, "non-finite initialization");
        }
        std::sort(recs.begin(), recs.end(), [](const Rec &a, const Rec &b) {
            const int32_t ka = total_key(a.dist), kb = total_key(b.dist);
            return ka != kb ? ka < kb : a.ordinal < b.ordinal;
        });
        double cumulative = 0.0;
        float median = 0.0f;
        for (const Rec &r : recs) {
            cumulative += static_cast<double>(r.w);
            if (cumulative >= total / 2.0) {
                median = r.dist;
                break;
            }
        }
        const double spread = median == 0.0f ? 1.0 : static_cast<double>(median);
        apply(static_cast<double>(scale_factor) / spread);
        return;
    }
    auto quantile = [](const std::vector<double> &s, double p) {
        const double h = static_cast<double>(s.size() - 1) * p;
        const size_t lo = static_cast<size_t>(std::floor(h)), hi = static_cast<size_t>(std::ceil(h));
        return s[lo] + (s[hi] - s[lo]) * (h - static_cast<double>(lo));
    };
    std::vector<double> spans(static_cast<size_t>(d)), scratch(static_cast<size_t>(n));
    for (int64_t c = 0; c < d; ++c) {
        for (int64_t r = 0; r < n; ++r) scratch[r] = v[r * d + c];
        std::sort(scratch.begin(), scratch.end());
        const double span = quantile(scratch, 0.95) - quantile(scratch, 0.05);
        spans[c] = span == 0.0 ? 1.0 : span;
    }
    const double target = std::log10(static_cast<double>(n)) * 3.0 *
                          std::log1p(static_cast<double>(repulsion)) / 0.6931471805599453 *
                          static_cast<double>(scale_factor);
    if (mode == 0) {
        double logs = 0.0;
        for (int64_t c = 0; c < d; ++c) logs += std::log(spans[c]);
        apply(target / std::exp(logs / static_cast<double>(d)));
    } else {
        for (int64_t c = 0; c < d; ++c) {
            const float mult = static_cast<float>(target / spans[c]);
            UMAPHIP_REQUIRE(finite(mult), UMAPHIP_NUMERICAL, "non-finite initialization");
            for (int64_t r = 0; r < n; ++r) {
                v[r * d + c] *= mult;
                UMAPHIP_REQUIRE(finite(v[r * d + c]), UMAPHIP_NUMERICAL,
                                "non-finite initialization");
            }
        }
    }
}

}  // namespace

// random_embedding (crates/umap-core/src/init.rs:1519-1530) then, for
// modes 0 and 1, the isotropic or per-axis rescale; mode -1 returns the raw
// U(-10, 10) draws (graph_edge callers rescale them with the graph).
umaphip_status umaphip_random_init(int64_t n, int64_t d, uint64_t seed, float repulsion,
                                   float scale_factor, int32_t mode, float *out) {
    return guarded([&] {
        UMAPHIP_REQUIRE(n > 0 && d > 0 && out && mode >= -1 && mode <= 1, UMAPHIP_INVALID_INPUT,
                        "invalid init arguments");
        if (n == 1) {
            std::fill(out, out + d, 0.0f);
            return;
        }
        StdRng rng(seed);
        for (int64_t i = 0; i < n * d; ++i) out[i] = rng.range_f32(-10.0f, 10.0f);
        if (mode >= 0) rescale(out, n, d, nullptr, nullptr, nullptr, mode, repulsion, scale_factor);
    });
}

umaphip_status umaphip_init_rescale(float *values, int64_t n, int64_t d, const int64_t *offsets,
                                    const int32_t *indices, const float *weights, int32_t mode,
                                    float repulsion, float scale_factor) {
    return guarded([&] {
        UMAPHIP_REQUIRE(n > 0 && d > 0 && values && mode >= 0 && mode <= 2,
                        UMAPHIP_INVALID_INPUT, "invalid rescale arguments");
        UMAPHIP_REQUIRE(mode != 2 || (offsets && (offsets[n] == 0 || (indices && weights))),
                        UMAPHIP_INVALID_INPUT, "graph_edge rescale needs the graph");
        if (mode == 2)
            for (int64_t h = 0; h < n; ++h)
                for (int64_t e = offsets[h]; e < offsets[h + 1]; ++e)
                    UMAPHIP_REQUIRE(indices[e] >= 0 && indices[e] < n, UMAPHIP_INVALID_INPUT,
                                    "graph_edge rescale: invalid edge endpoint");
        rescale(values, n, d, offsets, indices, weights, mode, repulsion, scale_factor);
    });
}

// weighted_initialization (crates/umap/src/standard_umap/estimator.rs:2385-
// 2481): rows with a coalesced membership of exactly 1 copy the smallest such
// neighbour; others take the slot-ordered sum of neighbour * normalised
// membership. normalised = directed / row sum (0 -> 1), summed in slot order.
umaphip_status umaphip_transform_init(const int32_t *indices, const float *directed, int64_t q,
                                      int64_t k, const float *embedding, int64_t n, int64_t d,
                                      float *out) {
    return guarded([&] {
        UMAPHIP_REQUIRE(indices && directed && embedding && out && q >= 0 && k > 0 && d > 0,
                        UMAPHIP_INVALID_INPUT, "invalid transform init arguments");
        std::atomic<int64_t> empty_row{-1};
        parallel_for(q, [&](int64_t row) {
            const int32_t *ids = indices + row * k;
            const float *w = directed + row * k;
            int64_t selected = -1;
            bool positive = false;
            for (int64_t s = 0; s < k; ++s) {
                const int32_t id = ids[s];
                if (id < 0) continue;
                bool seen = false;
                for (int64_t o = 0; o < s; ++o) seen = seen || ids[o] == id;
                if (seen) continue;
                float weight = 0.0f;
                for (int64_t o = 0; o < k; ++o)
                    if (ids[o] == id) weight += w[o];
                positive = positive || weight > 0.0f;
                if (weight == 1.0f && (selected < 0 || id < selected)) selected = id;
            }
            if (!positive) {
                int64_t expect = -1;
                empty_row.compare_exchange_strong(expect, row);
                return;
            }
            float *o = out + row * d;
            if (selected >= 0) {
                for (int64_t c = 0; c < d; ++c) o[c] = embedding[selected * d + c];
                return;
            }
            float sum = 0.0f;
            for (int64_t s = 0; s < k; ++s) sum += w[s];
            const float div = sum == 0.0f ? 1.0f : sum;
            for (int64_t c = 0; c < d; ++c) o[c] = 0.0f;
            for (int64_t s = 0; s < k; ++s) {
                const float nw = w[s] / div;
                const int64_t id = ids[s];
                for (int64_t c = 0; c < d; ++c) {
                    const float prod = embedding[id * d + c] * nw;
                    o[c] = o[c] + prod;
                }
            }
        }, 256);
        UMAPHIP_REQUIRE(empty_row.load() < 0, UMAPHIP_NUMERICAL,
                        "transform initialization row has no positive memberships");
    });
}

}  // extern "C"
