// Recursive (multilevel) initialisation, init="recursive" / "auto"
// (crates/umap-core/src/recursive_init/mod.rs, graph.rs, geometry.rs;
// driver crates/umap/src/standard_umap/estimator.rs:2506-2589).
//
// The fit graph is coarsened level by level on the host: hubs are chosen,
// labels spread by Jacobi voting, leftover rows take a nearby label, parts
// are aggregated into a summed graph whose row-normalised fuzzy union is the
// next level's graph. The coarsest level is placed from a sampled PCA
// anchor; every coarser layout runs on the device through the same layout
// engine as the main fit (umaphip_optimize_compact), then is prolongated to
// the finer level and rotated onto the anchor. The graph steps are
// O(edges log degree) host work spread over threads; the layouts are the
// only device work and dominate the time for large graphs.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <optional>
#include <vector>

#include "common.hpp"
#include "host_parallel.hpp"
#include "stdrng.hpp"

namespace umaphip {
namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// A square CSR graph (HostCsr, recursive_init/graph.rs:23-29) viewed
// through caller or owned arrays.
struct CsrView {
    int64_t n;
    const int64_t *ptr;
    const int32_t *idx;
    const float *w;
    int64_t edges() const { return ptr[n]; }
};

struct Csr {
    int64_t n = 0;
    std::vector<int64_t> ptr;
    std::vector<int32_t> idx;
    std::vector<float> w;
    CsrView view() const { return {n, ptr.data(), idx.data(), w.data()}; }
};

// SummedCsr (graph.rs:123-129): aggregated weights in f64, self loops kept.
struct Summed {
    int64_t n = 0;
    std::vector<int64_t> ptr;
    std::vector<int32_t> idx;
    std::vector<double> w;
};

struct Anchor {
    std::vector<float> values;  // [n, d]; rows outside the mask are zero
    std::vector<uint8_t> mask;  // sampled rows
    int64_t d = 0;
};

// SplitMix64 finaliser (graph.rs:133-138).
uint64_t mix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

// stream_seed (mod.rs:486-491): one independent seed per (domain, depth).
uint64_t stream_seed(uint64_t seed, uint64_t domain, uint64_t depth) {
    uint64_t v = seed ^ (depth * 0x9e3779b97f4a7c15ULL) ^ domain;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

int32_t total_key(float x) {
    int32_t b;
    std::memcpy(&b, &x, 4);
    return b ^ static_cast<int32_t>(static_cast<uint32_t>(b >> 31) >> 1);
}

// ------------------------------------------------------------ validation

// RecursiveInitConfig::validate (crates/umap-core/src/config.rs:404-430).
void check_params(const umaphip_recursive_params *p) {
    UMAPHIP_REQUIRE(p, UMAPHIP_INVALID_CONFIG, "missing recursive parameters");
    UMAPHIP_REQUIRE(p->coarsening_ratio >= 2 && p->coarsening_ratio <= 4, UMAPHIP_INVALID_CONFIG,
                    "coarsening_ratio must be 2, 3 or 4");
    UMAPHIP_REQUIRE(p->base_init_threshold >= 2 && p->base_init_threshold <= INT32_MAX,
                    UMAPHIP_INVALID_CONFIG,
                    "base_init_threshold must be at least 2 and fit in i32");
    UMAPHIP_REQUIRE(p->n_embedding_epochs >= 1, UMAPHIP_INVALID_CONFIG,
                    "n_embedding_epochs must be at least 1");
    UMAPHIP_REQUIRE(p->anchor_sample_size >= 1 && p->anchor_sample_size <= INT32_MAX,
                    UMAPHIP_INVALID_CONFIG,
                    "anchor_sample_size must be at least 1 and fit in i32");
    UMAPHIP_REQUIRE(p->hub_selection >= 0 && p->hub_selection <= 2, UMAPHIP_INVALID_CONFIG,
                    "hub_selection must be degree (0), diverse (1) or random (2)");
}

// Shape checks every entry point applies to a caller CSR.
void check_shape(const CsrView &g) {
    UMAPHIP_REQUIRE(g.n >= 0 && g.n < INT32_MAX && g.ptr && (g.n == 0 || g.ptr[0] == 0),
                    UMAPHIP_INVALID_INPUT, "invalid recursive CSR graph");
    for (int64_t r = 0; r < g.n; ++r)
        UMAPHIP_REQUIRE(g.ptr[r] <= g.ptr[r + 1], UMAPHIP_INVALID_INPUT,
                        "invalid recursive CSR graph");
    UMAPHIP_REQUIRE(g.edges() == 0 || (g.idx && g.w), UMAPHIP_INVALID_INPUT,
                    "invalid recursive CSR graph");
    for (int64_t e = 0; e < g.edges(); ++e)
        UMAPHIP_REQUIRE(g.idx[e] >= 0 && g.idx[e] < g.n, UMAPHIP_INVALID_INPUT,
                        "invalid recursive CSR graph");
}

// HostCsr::validate (graph.rs:31-69): sorted distinct columns, no self
// loops, finite positive weights, and symmetric support.
void validate_graph(const CsrView &g) {
    check_shape(g);
    std::atomic<int> bad{0};
    parallel_for(g.n, [&](int64_t r) {
        for (int64_t e = g.ptr[r]; e < g.ptr[r + 1]; ++e) {
            const int32_t c = g.idx[e];
            if ((e > g.ptr[r] && g.idx[e - 1] >= c) || c == r || !std::isfinite(g.w[e]) ||
                g.w[e] <= 0.0f) {
                bad.store(1);
                return;
            }
            const int32_t *b = g.idx + g.ptr[c], *en = g.idx + g.ptr[c + 1];
            if (!std::binary_search(b, en, static_cast<int32_t>(r))) bad.store(1);
        }
    }, 2048);
    UMAPHIP_REQUIRE(!bad.load(), UMAPHIP_INVALID_INPUT, "invalid recursive CSR graph");
}

// ------------------------------------------------------------ level params

struct Level {
    int64_t parts;
    int64_t epochs;
    float a, b;
    bool good;
    double curve_a, curve_b;
};

// LevelParams::new (mod.rs:541-560): the curve is taken to the 1/4 power at
// every level; good_initialization when the next level would coarsen again.
Level make_level(int64_t n, int64_t target, int64_t epochs, double ca, double cb,
                 const umaphip_recursive_params &p) {
    UMAPHIP_REQUIRE(epochs <= UINT32_MAX, UMAPHIP_OVERFLOW, "recursive layout epochs");
    Level L;
    L.parts = std::max<int64_t>(std::min<int64_t>(target, n / 2), 2);
    L.epochs = epochs;
    L.curve_a = std::pow(ca, 0.25);
    L.curve_b = std::pow(cb, 0.25);
    L.a = static_cast<float>(L.curve_a);
    L.b = static_cast<float>(L.curve_b);
    L.good = L.parts / p.coarsening_ratio > p.base_init_threshold;
    return L;
}

// LevelParams::first (mod.rs:510-522).
std::optional<Level> first_level(int64_t n, float a, float b, const umaphip_recursive_params &p) {
    if (n <= p.base_init_threshold) return std::nullopt;
    const int64_t target = std::max<int64_t>(n / p.coarsening_ratio, p.base_init_threshold);
    return make_level(n, target, p.n_embedding_epochs, static_cast<double>(a),
                      static_cast<double>(b), p);
}

// LevelParams::next (mod.rs:524-539): the next target comes from the
// requested parts; the epochs grow by sqrt(2) per level.
std::optional<Level> next_level(const Level &prev, int64_t n, const umaphip_recursive_params &p) {
    if (n <= p.base_init_threshold) return std::nullopt;
    const double epochs = std::floor(static_cast<double>(prev.epochs) * 1.4142135623730951);
    UMAPHIP_REQUIRE(epochs <= static_cast<double>(UINT32_MAX), UMAPHIP_OVERFLOW,
                    "recursive layout epochs");
    return make_level(n, prev.parts / p.coarsening_ratio, static_cast<int64_t>(epochs),
                      prev.curve_a, prev.curve_b, p);
}

// coarse_negative_range (mod.rs:563-567): half the rows, rounded to even.
int64_t coarse_negative_range(int64_t rows) {
    const int64_t half = rows / 2;
    return std::min<int64_t>(std::max<int64_t>(half + (rows % 2) * (half % 2), 1), rows);
}

// ------------------------------------------------------------ graph steps

// initialize_labels (graph.rs:140-197): hubs by weighted degree (ties by
// row), by degree with neighbours of earlier hubs blocked (diverse), or at
// hashed rows (random).
std::vector<int32_t> initialize_labels(const CsrView &g, int64_t parts, int method,
                                       uint64_t seed) {
    const int64_t n = g.n;
    UMAPHIP_REQUIRE(parts <= n && parts <= INT32_MAX, UMAPHIP_INVALID_INPUT,
                    "more recursive parts than rows");
    std::vector<int32_t> labels(static_cast<size_t>(n), -1);
    if (parts == 0) return labels;
    if (method == 2) {
        for (int64_t label = 0; label < parts; ++label)
            labels[mix(seed + static_cast<uint64_t>(label)) % static_cast<uint64_t>(n)] =
                static_cast<int32_t>(label);
        return labels;
    }
    // Rows by (degree total order, row): a stable LSD radix sort of the
    // unsigned order keys, starting from row order.
    std::vector<uint32_t> key(static_cast<size_t>(n));
    parallel_for(n, [&](int64_t r) {
        float sum = 0.0f;
        for (int64_t e = g.ptr[r]; e < g.ptr[r + 1]; ++e) sum += g.w[e];
        key[r] = static_cast<uint32_t>(total_key(sum)) ^ 0x80000000u;
    }, 4096);
    std::vector<int32_t> order(static_cast<size_t>(n)), tmp(static_cast<size_t>(n));
    std::iota(order.begin(), order.end(), 0);
    for (int shift = 0; shift < 32; shift += 11) {
        std::vector<int64_t> bucket(2049, 0);
        for (int64_t i = 0; i < n; ++i) ++bucket[((key[order[i]] >> shift) & 2047u) + 1];
        for (int b = 0; b < 2048; ++b) bucket[b + 1] += bucket[b];
        for (int64_t i = 0; i < n; ++i) tmp[bucket[(key[order[i]] >> shift) & 2047u]++] = order[i];
        order.swap(tmp);
    }
    if (method == 0) {
        for (int64_t label = 0; label < parts; ++label)
            labels[order[n - parts + label]] = static_cast<int32_t>(label);
        return labels;
    }
    std::vector<uint8_t> blocked(static_cast<size_t>(n), 0);
    int64_t next = 0;
    for (int64_t i = n - 1; i >= 0; --i) {
        const int32_t row = order[i];
        if (blocked[row]) continue;
        labels[row] = static_cast<int32_t>(next++);
        blocked[row] = 1;
        for (int64_t e = g.ptr[row]; e < g.ptr[row + 1]; ++e) blocked[g.idx[e]] = 1;
        if (next == parts) return labels;
    }
    for (int64_t i = n - 1; i >= 0; --i) {
        const int32_t row = order[i];
        if (labels[row] < 0) {
            labels[row] = static_cast<int32_t>(next++);
            if (next == parts) break;
        }
    }
    return labels;
}

// propagate_labels (graph.rs:199-253): Jacobi sweeps; an unlabelled row
// takes the label with the largest f32 vote (votes summed in edge order per
// label, labels visited by first edge, later ties win) once the vote reaches
// 1.0, or 10% of the row's total weight at deeper levels (share). Stops when
// every row is labelled or the unlabelled count stalls (checked every fifth
// sweep), at most 100 sweeps. Returns the sweeps run.
int propagate_labels(const CsrView &g, std::vector<int32_t> &labels, bool share) {
    const int64_t n = g.n;
    // Only unlabelled rows change, so each sweep visits the rows still
    // unlabelled (in row order) and copies back just those.
    std::vector<int32_t> active;
    for (int64_t r = 0; r < n; ++r)
        if (labels[r] < 0) active.push_back(static_cast<int32_t>(r));
    int64_t previous = static_cast<int64_t>(active.size());
    int32_t top = -1;
    for (int32_t l : labels) top = std::max(top, l);
    const size_t label_slots = static_cast<size_t>(top) + 1;
    std::vector<int32_t> next(labels);
    // Per row, votes are grouped by label with a stamped slot table: a label
    // opens its group at its first edge, so groups come out in first-edge
    // order and each group's f32 sum adds its votes in edge order, as the
    // reference's sort by (label, edge) and then by first edge does.
    struct Group {
        int32_t label;
        float sum;
    };
    static std::atomic<uint64_t> generation{0};
    for (int sweep = 0; sweep < 100; ++sweep) {
        const uint64_t tag = (generation.fetch_add(1) + 1) << 32;
        parallel_for(static_cast<int64_t>(active.size()), [&](int64_t a) {
            const int64_t r = active[a];
            thread_local std::vector<uint64_t> stamp;
            thread_local std::vector<int32_t> slot;
            thread_local std::vector<Group> groups;
            if (stamp.size() < label_slots) {
                stamp.resize(label_slots, 0);
                slot.resize(label_slots, 0);
            }
            const uint64_t mark = tag | static_cast<uint64_t>(r);
            groups.clear();
            double total = 0.0;
            for (int64_t e = g.ptr[r]; e < g.ptr[r + 1]; ++e) {
                const int32_t label = labels[g.idx[e]];
                total += static_cast<double>(g.w[e]);
                if (label < 0) continue;
                if (stamp[label] != mark) {
                    stamp[label] = mark;
                    slot[label] = static_cast<int32_t>(groups.size());
                    groups.push_back({label, g.w[e]});
                } else {
                    groups[slot[label]].sum += g.w[e];
                }
            }
            double maximum = share ? 0.1 * total : 1.0;
            for (const Group &grp : groups)
                if (static_cast<double>(grp.sum) >= maximum) {
                    maximum = static_cast<double>(grp.sum);
                    next[r] = grp.label;
                }
        }, 1024);
        size_t kept = 0;
        for (const int32_t r : active) {
            labels[r] = next[r];
            if (labels[r] < 0) active[kept++] = r;
        }
        active.resize(kept);
        if (sweep % 5 == 0) {
            const int64_t remaining = static_cast<int64_t>(active.size());
            if (remaining == 0 || remaining == previous) return sweep + 1;
            previous = remaining;
        }
    }
    return 100;
}

// label_outliers (graph.rs:255-297): rows in order; a depth-first walk of at
// most 100 pops takes the first labelled neighbour it meets, else a hashed
// label below the current label count. Labels found earlier are visible to
// later rows.
void label_outliers(const CsrView &g, std::vector<int32_t> &labels, uint64_t seed) {
    int32_t top = -1;
    for (int32_t l : labels) top = std::max(top, l);
    const uint64_t count = static_cast<uint64_t>(std::max<int64_t>(int64_t(top) + 1, 1));
    std::vector<int32_t> stack;
    for (int64_t row = 0; row < g.n; ++row) {
        if (labels[row] >= 0) continue;
        stack.clear();
        stack.push_back(static_cast<int32_t>(row));
        for (int step = 0; step < 100 && !stack.empty(); ++step) {
            const int32_t node = stack.back();
            stack.pop_back();
            for (int64_t e = g.ptr[node]; e < g.ptr[node + 1]; ++e) {
                const int32_t nb = g.idx[e];
                if (labels[nb] >= 0) {
                    labels[row] = labels[nb];
                    break;
                }
                stack.push_back(nb);
            }
            if (labels[row] >= 0) break;
        }
        if (labels[row] < 0)
            labels[row] = static_cast<int32_t>(mix(seed + static_cast<uint64_t>(row)) % count);
    }
}

// remap_labels (graph.rs:299-314): used labels to 0.. in increasing order,
// unlabelled rows after them in row order. Returns the part count.
// The rank of a label among the used labels is a prefix count over a
// presence table, which equals the reference's sort, dedup and binary search.
int64_t remap_labels(std::vector<int32_t> &labels) {
    int32_t top = -1;
    for (int32_t l : labels) top = std::max(top, l);
    std::vector<int32_t> rank(static_cast<size_t>(top) + 1, 0);
    for (int32_t l : labels)
        if (l >= 0) rank[l] = 1;
    int64_t next = 0;
    for (int32_t &r : rank) {
        const int32_t used = r;
        r = static_cast<int32_t>(next);
        next += used;
    }
    for (int32_t &l : labels) l = l < 0 ? static_cast<int32_t>(next++) : rank[l];
    return next;
}

// Rows grouped by part in row order (grouped_members, graph.rs:316-338).
void grouped_members(const std::vector<int32_t> &labels, int64_t parts,
                     std::vector<int64_t> &starts, std::vector<int32_t> &members) {
    starts.assign(static_cast<size_t>(parts + 1), 0);
    for (int32_t l : labels) ++starts[l + 1];
    for (int64_t i = 0; i < parts; ++i) starts[i + 1] += starts[i];
    std::vector<int64_t> cursor(starts.begin(), starts.end() - 1);
    members.assign(labels.size(), 0);
    for (size_t r = 0; r < labels.size(); ++r) members[cursor[labels[r]]++] = static_cast<int32_t>(r);
}

// Edges of the level being aggregated: the level-one f32 graph, or the
// previous level's summed graph.
struct AggInput {
    const int64_t *ptr;
    const int32_t *idx;
    const float *wf;
    const double *wd;
    double weight(int64_t e) const { return wd ? wd[e] : static_cast<double>(wf[e]); }
};

// aggregate_graph_counted (graph.rs:390-460): per part, the member rows'
// edges by target part, summed in f64 in edge order; self loops (edges
// inside a part) are kept. Members are visited in row order and each row's
// edges in order, so edge order is visiting order: a sparse accumulator
// (slot per target part, reset by a per-part stamp) adds each target's
// weights in exactly the order of the reference's (target, edge) sort, and
// only the distinct targets are sorted.
Summed aggregate(int64_t n, const AggInput &in, const std::vector<int32_t> &labels,
                 int64_t parts) {
    (void)n;
    std::vector<int64_t> starts;
    std::vector<int32_t> members;
    grouped_members(labels, parts, starts, members);
    // Stamps carry the pass's generation in their high half, so slots left
    // by an earlier pass or call on the same thread never match.
    static std::atomic<uint64_t> generation{0};
    thread_local std::vector<uint64_t> stamp;
    thread_local std::vector<int64_t> slot;
    auto prepare = [&] {
        if (stamp.size() < static_cast<size_t>(parts)) {
            stamp.resize(static_cast<size_t>(parts), 0);
            slot.resize(static_cast<size_t>(parts), 0);
        }
    };
    Summed s;
    s.n = parts;
    s.ptr.assign(static_cast<size_t>(parts + 1), 0);
    // Pass 1: distinct target parts per part (count_coarse_edges).
    const uint64_t count_tag = (generation.fetch_add(1) + 1) << 32;
    parallel_for(parts, [&](int64_t part) {
        prepare();
        const uint64_t mark = count_tag | static_cast<uint64_t>(part);
        int64_t c = 0;
        for (int64_t m = starts[part]; m < starts[part + 1]; ++m) {
            const int32_t r = members[m];
            for (int64_t e = in.ptr[r]; e < in.ptr[r + 1]; ++e) {
                const int32_t t = labels[in.idx[e]];
                if (stamp[t] != mark) {
                    stamp[t] = mark;
                    ++c;
                }
            }
        }
        s.ptr[part + 1] = c;
    }, 256);
    for (int64_t p = 0; p < parts; ++p) s.ptr[p + 1] += s.ptr[p];
    s.idx.resize(static_cast<size_t>(s.ptr[parts]));
    s.w.resize(static_cast<size_t>(s.ptr[parts]));
    // Pass 2: sums in edge order, then the part's targets sorted.
    const uint64_t fill_tag = (generation.fetch_add(1) + 1) << 32;
    parallel_for(parts, [&](int64_t part) {
        prepare();
        thread_local std::vector<std::pair<int32_t, double>> row;
        const uint64_t mark = fill_tag | static_cast<uint64_t>(part);
        int64_t o = s.ptr[part];
        for (int64_t m = starts[part]; m < starts[part + 1]; ++m) {
            const int32_t r = members[m];
            for (int64_t e = in.ptr[r]; e < in.ptr[r + 1]; ++e) {
                const int32_t t = labels[in.idx[e]];
                const double w = in.weight(e);
                if (stamp[t] != mark) {
                    stamp[t] = mark;
                    slot[t] = o;
                    s.idx[o] = t;
                    s.w[o++] = w;
                } else {
                    s.w[slot[t]] += w;
                }
            }
        }
        const int64_t b = s.ptr[part], en = s.ptr[part + 1];
        row.clear();
        for (int64_t q = b; q < en; ++q) row.push_back({s.idx[q], s.w[q]});
        std::sort(row.begin(), row.end(),
                  [](const std::pair<int32_t, double> &x, const std::pair<int32_t, double> &y) {
                      return x.first < y.first;
                  });
        for (int64_t q = b; q < en; ++q) {
            s.idx[q] = row[q - b].first;
            s.w[q] = row[q - b].second;
        }
    }, 256);
    return s;
}

// transition_graph (graph.rs:462-533): drop self loops, normalise each row
// by its off-diagonal total, and take the fuzzy union f + r - f r (clamped
// to [0, 1], as f32) over the symmetric support; zero weights are dropped.
// Built per row by merging the row's own edges with its incoming ones, which
// gives the sorted, de-duplicated support of the reference's global sort.
Csr transition(const Summed &s) {
    const int64_t n = s.n;
    std::vector<double> totals(static_cast<size_t>(n), 0.0);
    parallel_for(n, [&](int64_t r) {
        double t = 0.0;
        for (int64_t e = s.ptr[r]; e < s.ptr[r + 1]; ++e)
            if (s.idx[e] != r) t += s.w[e];
        totals[r] = t;
    }, 4096);
    auto valid = [&](int64_t r, int64_t e) {
        return totals[r] > 0.0 && s.idx[e] != r && s.w[e] != 0.0;
    };
    // Incoming valid edges, by target, in source order: a counting sort
    // over contiguous row chunks (chunk c's entries for a target precede
    // chunk c + 1's), each chunk counting and placing its own edges.
    const int64_t chunks =
        std::min<int64_t>(host_workers(), std::max<int64_t>(1, s.ptr[n] / (1 << 16)));
    std::vector<int64_t> cnt(static_cast<size_t>(chunks * n), 0);
    parallel_blocks(n, chunks, [&](int64_t c, int64_t b, int64_t en) {
        int64_t *cc = &cnt[c * n];
        for (int64_t r = b; r < en; ++r)
            for (int64_t e = s.ptr[r]; e < s.ptr[r + 1]; ++e)
                if (valid(r, e)) ++cc[s.idx[e]];
    });
    std::vector<int64_t> tptr(static_cast<size_t>(n + 1), 0);
    for (int64_t t = 0, run = 0; t < n; ++t) {
        for (int64_t c = 0; c < chunks; ++c) {
            const int64_t k = cnt[c * n + t];
            cnt[c * n + t] = run;
            run += k;
        }
        tptr[t + 1] = run;
    }
    std::vector<int32_t> tsrc(static_cast<size_t>(tptr[n]));
    std::vector<double> tw(static_cast<size_t>(tptr[n]));
    parallel_blocks(n, chunks, [&](int64_t c, int64_t b, int64_t en) {
        int64_t *cc = &cnt[c * n];
        for (int64_t r = b; r < en; ++r)
            for (int64_t e = s.ptr[r]; e < s.ptr[r + 1]; ++e)
                if (valid(r, e)) {
                    const int64_t o = cc[s.idx[e]]++;
                    tsrc[o] = static_cast<int32_t>(r);
                    tw[o] = s.w[e] / totals[r];
                }
    });
    // Walk the merged support of row r: fn(col, forward, reverse).
    auto walk = [&](int64_t r, auto &&fn) {
        int64_t e = s.ptr[r], t = tptr[r];
        const int64_t ee = s.ptr[r + 1], te = tptr[r + 1];
        for (;;) {
            while (e < ee && !valid(r, e)) ++e;
            if (e >= ee && t >= te) break;
            const int32_t ce = e < ee ? s.idx[e] : INT32_MAX;
            const int32_t ct = t < te ? tsrc[t] : INT32_MAX;
            if (ce == ct) {
                fn(ce, s.w[e] / totals[r], tw[t]);
                ++e;
                ++t;
            } else if (ce < ct) {
                // reverse edge absent or invalid: directed(c, r) is zero
                // unless c -> r exists with a positive total; it would then
                // be in the incoming list, so it is zero here.
                fn(ce, s.w[e] / totals[r], 0.0);
                ++e;
            } else {
                fn(ct, 0.0, tw[t]);
                ++t;
            }
        }
    };
    auto weight = [](double f, double r) {
        return static_cast<float>(std::min(std::max(f + r - f * r, 0.0), 1.0));
    };
    Csr out;
    out.n = n;
    out.ptr.assign(static_cast<size_t>(n + 1), 0);
    parallel_for(n, [&](int64_t r) {
        int64_t c = 0;
        walk(r, [&](int32_t, double f, double rv) { c += weight(f, rv) != 0.0f; });
        out.ptr[r + 1] = c;
    }, 2048);
    for (int64_t r = 0; r < n; ++r) out.ptr[r + 1] += out.ptr[r];
    out.idx.resize(static_cast<size_t>(out.ptr[n]));
    out.w.resize(static_cast<size_t>(out.ptr[n]));
    parallel_for(n, [&](int64_t r) {
        int64_t o = out.ptr[r];
        walk(r, [&](int32_t col, double f, double rv) {
            const float w = weight(f, rv);
            if (w != 0.0f) {
                out.idx[o] = col;
                out.w[o++] = w;
            }
        });
    }, 2048);
    return out;
}

// ------------------------------------------------------------ geometry

// Row-major f64 matrix.
struct Mat {
    int64_t rows = 0, cols = 0;
    std::vector<double> a;
    Mat() = default;
    Mat(int64_t r, int64_t c) : rows(r), cols(c), a(static_cast<size_t>(r * c), 0.0) {}
    double &operator()(int64_t i, int64_t j) { return a[i * cols + j]; }
    double operator()(int64_t i, int64_t j) const { return a[i * cols + j]; }
};

// C = A B, rows of A split over threads.
Mat product(const Mat &A, const Mat &B) {
    Mat C(A.rows, B.cols);
    parallel_for(A.rows, [&](int64_t i) {
        double *c = &C.a[i * C.cols];
        for (int64_t k = 0; k < A.cols; ++k) {
            const double x = A(i, k);
            const double *b = &B.a[k * B.cols];
            for (int64_t j = 0; j < B.cols; ++j) c[j] += x * b[j];
        }
    }, 64);
    return C;
}

// C = A^T B for tall A and B (contraction over their rows): per-thread
// partial products reduced in block order.
Mat product_tn(const Mat &A, const Mat &B) {
    const int64_t workers = std::min<int64_t>(host_workers(), std::max<int64_t>(1, A.rows / 256));
    std::vector<Mat> part(static_cast<size_t>(workers), Mat(A.cols, B.cols));
    parallel_blocks(A.rows, workers, [&](int64_t w, int64_t b, int64_t e) {
        Mat &C = part[w];
        for (int64_t i = b; i < e; ++i)
            for (int64_t k = 0; k < A.cols; ++k) {
                const double x = A(i, k);
                double *c = &C.a[k * C.cols];
                const double *bb = &B.a[i * B.cols];
                for (int64_t j = 0; j < B.cols; ++j) c[j] += x * bb[j];
            }
    });
    for (int64_t w = 1; w < workers; ++w)
        for (size_t i = 0; i < part[0].a.size(); ++i) part[0].a[i] += part[w].a[i];
    return std::move(part[0]);
}

// Thin Q (m x min(m, n)) of a Householder QR of Y.
Mat thin_q(Mat Y) {
    const int64_t m = Y.rows, n = Y.cols, k = std::min(m, n);
    std::vector<std::vector<double>> vs;
    std::vector<double> taus;
    for (int64_t j = 0; j < k; ++j) {
        double norm = 0.0;
        for (int64_t i = j; i < m; ++i) norm += Y(i, j) * Y(i, j);
        norm = std::sqrt(norm);
        std::vector<double> v(static_cast<size_t>(m - j), 0.0);
        double tau = 0.0;
        if (norm > 0.0) {
            const double alpha = Y(j, j) > 0.0 ? -norm : norm;
            for (int64_t i = j; i < m; ++i) v[i - j] = Y(i, j);
            v[0] -= alpha;
            double vv = 0.0;
            for (double x : v) vv += x * x;
            if (vv > 0.0) {
                tau = 2.0 / vv;
                for (int64_t c = j; c < n; ++c) {
                    double dot = 0.0;
                    for (int64_t i = j; i < m; ++i) dot += v[i - j] * Y(i, c);
                    dot *= tau;
                    for (int64_t i = j; i < m; ++i) Y(i, c) -= dot * v[i - j];
                }
            }
        }
        vs.push_back(std::move(v));
        taus.push_back(tau);
    }
    Mat Q(m, k);
    for (int64_t i = 0; i < k; ++i) Q(i, i) = 1.0;
    for (int64_t j = k - 1; j >= 0; --j) {
        const auto &v = vs[j];
        if (taus[j] == 0.0) continue;
        for (int64_t c = 0; c < k; ++c) {
            double dot = 0.0;
            for (int64_t i = j; i < m; ++i) dot += v[i - j] * Q(i, c);
            dot *= taus[j];
            for (int64_t i = j; i < m; ++i) Q(i, c) -= dot * v[i - j];
        }
    }
    return Q;
}

// Thin SVD B = U diag(s) V^T with singular values descending: one-sided
// (Hestenes) Jacobi on the columns of B^T. U is rows(B) x k, V is
// cols(B) x k, k = min(rows, cols). Left vectors of zero singular values
// are completed to an orthonormal set.
void thin_svd(const Mat &B, Mat &U, std::vector<double> &s, Mat &V) {
    const int64_t m = B.rows, p = B.cols, k = std::min(m, p);
    // Columns of A = B^T (length p), one per row of B, stored contiguously.
    std::vector<double> A(static_cast<size_t>(m * p));
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < p; ++j) A[i * p + j] = B(i, j);
    std::vector<double> W(static_cast<size_t>(m * m), 0.0);  // rotations, m x m
    for (int64_t i = 0; i < m; ++i) W[i * m + i] = 1.0;
    const double eps = std::numeric_limits<double>::epsilon();
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int64_t x = 0; x < m - 1; ++x)
            for (int64_t y = x + 1; y < m; ++y) {
                double a = 0.0, b = 0.0, c = 0.0;
                const double *cx = &A[x * p], *cy = &A[y * p];
                for (int64_t t = 0; t < p; ++t) {
                    a += cx[t] * cx[t];
                    b += cy[t] * cy[t];
                    c += cx[t] * cy[t];
                }
                if (c == 0.0 || std::fabs(c) <= eps * std::sqrt(a * b)) continue;
                rotated = true;
                const double zeta = (b - a) / (2.0 * c);
                const double tn = (zeta >= 0.0 ? 1.0 : -1.0) /
                                  (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double cs = 1.0 / std::sqrt(1.0 + tn * tn), sn = cs * tn;
                double *px = &A[x * p], *py = &A[y * p];
                for (int64_t t = 0; t < p; ++t) {
                    const double u = px[t], v = py[t];
                    px[t] = cs * u - sn * v;
                    py[t] = sn * u + cs * v;
                }
                for (int64_t t = 0; t < m; ++t) {
                    const double u = W[t * m + x], v = W[t * m + y];
                    W[t * m + x] = cs * u - sn * v;
                    W[t * m + y] = sn * u + cs * v;
                }
            }
        if (!rotated) break;
    }
    std::vector<double> norms(static_cast<size_t>(m));
    for (int64_t i = 0; i < m; ++i) {
        double q = 0.0;
        for (int64_t t = 0; t < p; ++t) q += A[i * p + t] * A[i * p + t];
        norms[i] = std::sqrt(q);
    }
    std::vector<int64_t> order(static_cast<size_t>(m));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t x, int64_t y) { return norms[x] > norms[y]; });
    U = Mat(m, k);
    V = Mat(p, k);
    s.assign(static_cast<size_t>(k), 0.0);
    for (int64_t c = 0; c < k; ++c) {
        const int64_t o = order[c];
        s[c] = norms[o];
        for (int64_t i = 0; i < m; ++i) U(i, c) = W[i * m + o];
        if (norms[o] > 0.0)
            for (int64_t t = 0; t < p; ++t) V(t, c) = A[o * p + t] / norms[o];
    }
    // Complete right vectors of zero singular values (U is orthonormal by
    // construction) with Gram-Schmidt against the unit basis.
    for (int64_t c = 0; c < k; ++c) {
        if (s[c] > 0.0) continue;
        for (int64_t basis = 0; basis < p; ++basis) {
            std::vector<double> v(static_cast<size_t>(p), 0.0);
            v[basis] = 1.0;
            for (int pass = 0; pass < 2; ++pass)
                for (int64_t o = 0; o < k; ++o) {
                    if (o == c || (s[o] == 0.0 && o > c)) continue;
                    double dot = 0.0;
                    for (int64_t t = 0; t < p; ++t) dot += V(t, o) * v[t];
                    for (int64_t t = 0; t < p; ++t) v[t] -= dot * V(t, o);
                }
            double q = 0.0;
            for (double x : v) q += x * x;
            if (q > 1e-12) {
                q = std::sqrt(q);
                for (int64_t t = 0; t < p; ++t) V(t, c) = v[t] / q;
                break;
            }
        }
    }
}

// anchor_reference (geometry.rs:266-427): reservoir-sample up to
// sample_size rows, centre them, take a randomised PCA (one power step,
// Gaussian test matrix of d + 8 columns), flip each axis so that its
// largest-magnitude sample score is positive (sklearn svd_flip), and map
// every axis of the sampled rows to [0, 10]. Axes beyond the rank are
// uniform in [0, 10).
Anchor anchor_reference(const float *data, int64_t n, int64_t features, int64_t d,
                        int64_t sample_size, StdRng &rng) {
    UMAPHIP_REQUIRE(n > 0 && features > 0 && d > 0 && sample_size > 0 && data,
                    UMAPHIP_INVALID_INPUT, "invalid recursive anchor shape");
    const int64_t s = std::min(n, sample_size);
    std::vector<int64_t> sample(static_cast<size_t>(s));
    std::iota(sample.begin(), sample.end(), 0);
    for (int64_t i = s; i < n; ++i) {
        UMAPHIP_REQUIRE(i < INT32_MAX, UMAPHIP_OVERFLOW, "anchor sample range");
        const int64_t j = rng.below(static_cast<uint32_t>(i + 1));
        if (j < s) sample[j] = i;
    }
    std::sort(sample.begin(), sample.end());
    Anchor A;
    A.d = d;
    A.mask.assign(static_cast<size_t>(n), 0);
    for (int64_t r : sample) A.mask[r] = 1;
    std::vector<double> means(static_cast<size_t>(features), 0.0);
    for (int64_t r : sample)
        for (int64_t j = 0; j < features; ++j) means[j] += static_cast<double>(data[r * features + j]);
    for (double &x : means) x /= static_cast<double>(s);
    const bool transposed = s < features;
    const int64_t m = transposed ? features : s, p = transposed ? s : features;
    Mat M(m, p);
    parallel_for(m, [&](int64_t i) {
        for (int64_t j = 0; j < p; ++j) {
            const int64_t row = transposed ? j : i, col = transposed ? i : j;
            M(i, j) = static_cast<double>(data[sample[row] * features + col]) - means[col];
        }
    }, 64);
    const int64_t width = d + 8;
    Mat omega(p, width);
    bool have_spare = false;
    double spare = 0.0;
    for (int64_t i = 0; i < p; ++i)
        for (int64_t j = 0; j < width; ++j) {
            if (have_spare) {
                omega(i, j) = spare;
                have_spare = false;
            } else {
                double x, y;
                rng.gaussian_pair(x, y);
                omega(i, j) = x;
                spare = y;
                have_spare = true;
            }
        }
    // Y = M (M^T (M Omega)), no normalisation between the products.
    const Mat first = product(M, omega);
    const Mat power = product_tn(M, first);
    const Mat Q = thin_q(product(M, power));
    const Mat B = product_tn(Q, M);  // Q^T M
    Mat U, V;
    std::vector<double> sv;
    thin_svd(B, U, sv, V);
    const int64_t rank = std::min(std::min(d, s), features);
    const Mat projected = product(Q, U);
    A.values.assign(static_cast<size_t>(n * d), 0.0f);
    for (int64_t axis = 0; axis < rank; ++axis) {
        auto coord = [&](int64_t i) { return transposed ? V(i, axis) : projected(i, axis); };
        const int64_t len = transposed ? s : m;
        int64_t pivot = 0;
        for (int64_t i = 1; i < len; ++i)
            if (std::fabs(coord(i)) > std::fabs(coord(pivot))) pivot = i;
        const double sign = coord(pivot) < 0.0 ? -1.0 : 1.0;
        const double sigma = sv[axis];
        for (int64_t i = 0; i < s; ++i) {
            const float v = static_cast<float>(sign * sigma * coord(i));
            UMAPHIP_REQUIRE(std::isfinite(v), UMAPHIP_NUMERICAL, "non-finite recursive PCA score");
            A.values[sample[i] * d + axis] = v;
        }
        float lo = INFINITY;
        for (int64_t r : sample) lo = std::min(lo, A.values[r * d + axis]);
        float span = 0.0f;
        for (int64_t r : sample) {
            float &x = A.values[r * d + axis];
            x -= lo;
            span = std::max(span, x);
        }
        UMAPHIP_REQUIRE(std::isfinite(span), UMAPHIP_NUMERICAL, "non-finite recursive anchor span");
        if (span == 0.0f) span = 1.0f;
        for (int64_t r : sample) {
            float &x = A.values[r * d + axis];
            x /= span;
            x *= 10.0f;
        }
    }
    for (int64_t r : sample)
        for (int64_t axis = rank; axis < d; ++axis) A.values[r * d + axis] = rng.range_f32(0.0f, 10.0f);
    return A;
}

// coarse_anchor (geometry.rs:429-458): f64 mean of each part's sampled rows.
Anchor coarse_anchor(const Anchor &A, const std::vector<int32_t> &labels, int64_t parts) {
    const int64_t d = A.d;
    std::vector<double> sums(static_cast<size_t>(parts * d), 0.0);
    std::vector<int64_t> counts(static_cast<size_t>(parts), 0);
    for (size_t r = 0; r < labels.size(); ++r) {
        if (!A.mask[r]) continue;
        const int64_t part = labels[r];
        ++counts[part];
        for (int64_t j = 0; j < d; ++j) sums[part * d + j] += static_cast<double>(A.values[r * d + j]);
    }
    Anchor C;
    C.d = d;
    C.values.assign(static_cast<size_t>(parts * d), 0.0f);
    C.mask.assign(static_cast<size_t>(parts), 0);
    for (int64_t p = 0; p < parts; ++p)
        if (counts[p] > 0) {
            C.mask[p] = 1;
            for (int64_t j = 0; j < d; ++j)
                C.values[p * d + j] = static_cast<float>(sums[p * d + j] / static_cast<double>(counts[p]));
        }
    return C;
}

// base_case (geometry.rs:460-539): sampled rows keep the anchor, others take
// the weighted mean of their sampled neighbours (the anchor mean without
// any); then each axis is centred and scaled so its 5-95% span is
// 3 log10(n) log2(depth + 1).
std::vector<float> base_case(const CsrView &g, const Anchor &A, int64_t depth) {
    const int64_t n = g.n, d = A.d;
    std::vector<double> res(static_cast<size_t>(n * d), 0.0), mean(static_cast<size_t>(d), 0.0);
    int64_t count = 0;
    for (int64_t r = 0; r < n; ++r) count += A.mask[r] != 0;
    UMAPHIP_REQUIRE(count > 0, UMAPHIP_INVALID_INPUT, "recursive anchor has no sampled rows");
    for (int64_t r = 0; r < n; ++r)
        if (A.mask[r])
            for (int64_t j = 0; j < d; ++j) mean[j] += static_cast<double>(A.values[r * d + j]);
    for (double &x : mean) x /= static_cast<double>(count);
    for (int64_t r = 0; r < n; ++r) {
        if (A.mask[r]) {
            for (int64_t j = 0; j < d; ++j) res[r * d + j] = static_cast<double>(A.values[r * d + j]);
            continue;
        }
        double total = 0.0;
        for (int64_t e = g.ptr[r]; e < g.ptr[r + 1]; ++e) {
            const int32_t nb = g.idx[e];
            if (!A.mask[nb]) continue;
            const double w = static_cast<double>(g.w[e]);
            total += w;
            for (int64_t j = 0; j < d; ++j) res[r * d + j] += w * static_cast<double>(A.values[nb * d + j]);
        }
        for (int64_t j = 0; j < d; ++j) res[r * d + j] = total > 0.0 ? res[r * d + j] / total : mean[j];
    }
    std::vector<double> sorted(static_cast<size_t>(n));
    const double scale = std::log10(static_cast<double>(n)) * 3.0 *
                         std::log2(static_cast<double>(depth + 1));
    for (int64_t j = 0; j < d; ++j) {
        double sum = 0.0;
        for (int64_t i = 0; i < n; ++i) sum += res[i * d + j];
        const double avg = sum / static_cast<double>(n);
        for (int64_t i = 0; i < n; ++i) {
            res[i * d + j] -= avg;
            sorted[i] = res[i * d + j];
        }
        std::sort(sorted.begin(), sorted.end());
        auto quantile = [&](double q) {
            const double pos = q * static_cast<double>(n - 1);
            const int64_t lo = static_cast<int64_t>(std::floor(pos)), hi = static_cast<int64_t>(std::ceil(pos));
            return sorted[lo] + (sorted[hi] - sorted[lo]) * (pos - static_cast<double>(lo));
        };
        const double spread = quantile(0.95) - quantile(0.05);
        const double mult = scale / (spread == 0.0 ? 1.0 : spread);
        for (int64_t i = 0; i < n; ++i) res[i * d + j] *= mult;
    }
    std::vector<float> out(res.size());
    for (size_t i = 0; i < res.size(); ++i) {
        out[i] = static_cast<float>(res[i]);
        UMAPHIP_REQUIRE(std::isfinite(out[i]), UMAPHIP_NUMERICAL, "non-finite recursive base layout");
    }
    return out;
}

// prolongate (geometry.rs:541-582): each row is the mean of its part's
// coordinates and the weighted mean of its neighbours' parts' coordinates,
// then the columns are centred.
std::vector<float> prolongate(const CsrView &g, const std::vector<int32_t> &labels,
                              const float *coarse, int64_t d) {
    const int64_t n = g.n;
    std::vector<double> res(static_cast<size_t>(n * d), 0.0);
    parallel_for(n, [&](int64_t r) {
        double total = 0.0;
        double *out = &res[r * d];
        for (int64_t e = g.ptr[r]; e < g.ptr[r + 1]; ++e) {
            const double w = static_cast<double>(g.w[e]);
            total += w;
            const int64_t p = labels[g.idx[e]];
            for (int64_t j = 0; j < d; ++j) out[j] += w * static_cast<double>(coarse[p * d + j]);
        }
        const int64_t own = labels[r];
        for (int64_t j = 0; j < d; ++j)
            out[j] = 0.5 * ((total > 0.0 ? out[j] / total : 0.0) +
                            static_cast<double>(coarse[own * d + j]));
    }, 2048);
    std::vector<float> out(res.size());
    for (int64_t j = 0; j < d; ++j) {
        double sum = 0.0;
        for (int64_t i = 0; i < n; ++i) sum += res[i * d + j];
        const double mean = sum / static_cast<double>(n);
        for (int64_t i = 0; i < n; ++i) {
            out[i * d + j] = static_cast<float>(res[i * d + j] - mean);
            UMAPHIP_REQUIRE(std::isfinite(out[i * d + j]), UMAPHIP_NUMERICAL,
                            "non-finite recursive prolongation");
        }
    }
    return out;
}

// procrustes_align (geometry.rs:584-690): the proper rotation (no
// reflection, no scaling) that best aligns the sampled rows' centred layout
// with their centred anchor, applied to every row about the origin.
void procrustes_align(float *layout, int64_t n, const Anchor &A) {
    const int64_t d = A.d;
    int64_t count = 0;
    for (int64_t i = 0; i < n; ++i) count += A.mask[i] != 0;
    UMAPHIP_REQUIRE(count > 0, UMAPHIP_INVALID_INPUT, "recursive anchor has no sampled rows");
    std::vector<double> am(static_cast<size_t>(d), 0.0), xm(static_cast<size_t>(d), 0.0);
    for (int64_t i = 0; i < n; ++i)
        if (A.mask[i])
            for (int64_t j = 0; j < d; ++j) {
                am[j] += static_cast<double>(A.values[i * d + j]);
                xm[j] += static_cast<double>(layout[i * d + j]);
            }
    for (int64_t j = 0; j < d; ++j) {
        am[j] /= static_cast<double>(count);
        xm[j] /= static_cast<double>(count);
    }
    double an = 0.0, xn = 0.0;
    for (int64_t i = 0; i < n; ++i)
        if (A.mask[i])
            for (int64_t j = 0; j < d; ++j) {
                const double u = static_cast<double>(A.values[i * d + j]) - am[j];
                const double v = static_cast<double>(layout[i * d + j]) - xm[j];
                an += u * u;
                xn += v * v;
            }
    if (an == 0.0 || xn == 0.0) return;
    const double scale = std::sqrt(an * xn);
    Mat cov(d, d);
    for (int64_t j = 0; j < d; ++j)
        for (int64_t k = 0; k < d; ++k) {
            double sum = 0.0;
            for (int64_t i = 0; i < n; ++i)
                if (A.mask[i])
                    sum += (static_cast<double>(layout[i * d + j]) - xm[j]) *
                           (static_cast<double>(A.values[i * d + k]) - am[k]);
            cov(j, k) = sum / scale;
        }
    Mat U, V;
    std::vector<double> sv;
    thin_svd(cov, U, sv, V);
    auto rotation_of = [&] {
        Mat R(d, d);
        for (int64_t i = 0; i < d; ++i)
            for (int64_t j = 0; j < d; ++j) {
                double sum = 0.0;
                for (int64_t k = 0; k < d; ++k) sum += U(i, k) * V(j, k);
                R(i, j) = sum;
            }
        return R;
    };
    Mat R = rotation_of();
    // Sign of det(R) by partial-pivot elimination.
    Mat w = R;
    double det = 1.0;
    for (int64_t j = 0; j < d; ++j) {
        int64_t pivot = j;
        for (int64_t i = j + 1; i < d; ++i)
            if (std::fabs(w(i, j)) > std::fabs(w(pivot, j))) pivot = i;
        if (pivot != j) {
            for (int64_t k = j; k < d; ++k) std::swap(w(j, k), w(pivot, k));
            det = -det;
        }
        const double diag = w(j, j);
        det *= diag;
        if (diag != 0.0)
            for (int64_t i = j + 1; i < d; ++i) {
                const double f = w(i, j) / diag;
                for (int64_t k = j + 1; k < d; ++k) w(i, k) -= f * w(j, k);
            }
    }
    if (det < 0.0) {
        for (int64_t i = 0; i < d; ++i) U(i, d - 1) = -U(i, d - 1);
        R = rotation_of();
    }
    std::atomic<int> bad{0};
    parallel_for(n, [&](int64_t i) {
        float row[64];
        std::vector<float> wide;
        float *out = row;
        if (d > 64) {
            wide.resize(static_cast<size_t>(d));
            out = wide.data();
        }
        for (int64_t j = 0; j < d; ++j) {
            double sum = 0.0;
            for (int64_t k = 0; k < d; ++k) sum += static_cast<double>(layout[i * d + k]) * R(k, j);
            out[j] = static_cast<float>(sum);
            if (!std::isfinite(out[j])) bad.store(1);
        }
        std::memcpy(layout + i * d, out, static_cast<size_t>(d) * sizeof(float));
    }, 4096);
    UMAPHIP_REQUIRE(!bad.load(), UMAPHIP_NUMERICAL, "non-finite recursive rotation");
}

// ------------------------------------------------------------ recursion

// edge_periods (crates/umap-core/src/modern_layout.rs:188-208); false for an
// invalid period (called from worker threads, so it does not throw).
bool edge_periods(float w, float wmax, int64_t epochs, float rate, double &period,
                  double &negative) {
    const float samples = static_cast<float>(epochs) * (w / wmax);
    period = static_cast<double>(epochs) / static_cast<double>(samples);
    negative = rate == 0.0f ? INFINITY : (period / static_cast<double>(rate)) * 1.5;
    return std::isfinite(period) && period > 0.0 &&
           (rate <= 0.0f || (std::isfinite(negative) && negative > 0.0));
}

struct Run {
    const umaphip_recursive_params &P;
    uint64_t seed;
    int64_t d;
    umaphip_recursive_stats *stats;
    double layout_ms = 0.0;

    // The coarse layout of one level (estimator.rs:2529-2585 with
    // ModernLayoutParams::recursive_level, crates/umap-core/src/
    // modern_layout.rs:55-71): compact_graph_from_csr (:210-300) with
    // negative rate 1, Adam, repulsion 4, initial alpha 0.5, one negative
    // sample per edge over half the realised rows, adapted scale.
    void layout(const Csr &g, std::vector<float> &values, const Level &L, uint64_t level_seed) {
        const auto t0 = Clock::now();
        const int64_t rows = g.n, E = g.ptr[rows];
        UMAPHIP_REQUIRE(rows < UINT32_MAX && E < UINT32_MAX, UMAPHIP_OVERFLOW,
                        "recursive coarse graph too large");
        std::vector<uint32_t> offsets(static_cast<size_t>(rows + 1)), indices(static_cast<size_t>(E));
        std::vector<double> periods(static_cast<size_t>(E)), negative(static_cast<size_t>(E));
        std::vector<double> next_pos(static_cast<size_t>(E)), next_neg(static_cast<size_t>(E));
        for (int64_t r = 0; r <= rows; ++r) offsets[r] = static_cast<uint32_t>(g.ptr[r]);
        float wmax = 0.0f;
        for (float w : g.w) wmax = std::max(wmax, w);
        std::atomic<int> bad{0};
        parallel_for(E, [&](int64_t e) {
            indices[e] = static_cast<uint32_t>(g.idx[e]);
            if (!edge_periods(g.w[e], wmax, L.epochs, 1.0f, periods[e], negative[e])) bad.store(1);
            next_pos[e] = periods[e];
            next_neg[e] = negative[e];
        }, 1 << 15);
        UMAPHIP_REQUIRE(!bad.load(), UMAPHIP_NUMERICAL, "invalid layout period");
        UMAPHIP_REQUIRE(L.epochs <= INT32_MAX, UMAPHIP_OVERFLOW, "recursive layout epochs");
        umaphip_layout_params lp{};
        lp.optimizer = 0;
        lp.epochs = static_cast<int32_t>(L.epochs);
        lp.a = L.a;
        lp.b = L.b;
        lp.repulsion = 4.0f;
        lp.initial_alpha = 0.5f;
        lp.negative_sample_rate = 1;
        lp.negative_selection_range = coarse_negative_range(rows);
        lp.has_scale = 0;
        lp.scale = 0.0f;
        lp.adaptation_samples = 128;
        lp.exclude_graph_neighbors = 0;
        lp.good_initialization = L.good ? 1 : 0;
        lp.seed = level_seed;
        const umaphip_status st = umaphip_optimize_compact(
            rows, rows, d, values.data(), nullptr, offsets.data(), indices.data(), E,
            periods.data(), next_pos.data(), negative.data(), next_neg.data(), &lp, 1, nullptr);
        if (st != UMAPHIP_OK) throw Fail{st};
        layout_ms += ms_since(t0);
    }

    // Adds the time since t to host stage `stage` and restarts t.
    void lap(int stage, Clock::time_point &t) {
        const auto now = Clock::now();
        if (stats) stats->stage_ms[stage] += std::chrono::duration<double, std::milli>(now - t).count();
        t = now;
    }

    void record(int64_t depth, int64_t rows, int64_t edges, int sweeps) {
        if (!stats || depth < 1 || depth > 32) return;
        stats->depth = std::max<int32_t>(stats->depth, static_cast<int32_t>(depth));
        stats->rows[depth - 1] = rows;
        stats->edges[depth - 1] = edges;
        stats->sweeps[depth - 1] = sweeps;
    }

    // descend (mod.rs:125-320).
    std::vector<float> descend(const CsrView &g, const Summed *summed, const Anchor &anchor,
                               std::optional<Level> params, int64_t depth) {
        auto t = Clock::now();
        if (!params) {
            record(depth, g.n, g.edges(), 0);
            std::vector<float> out = base_case(g, anchor, depth);
            lap(7, t);
            return out;
        }
        const Level L = *params;
        std::vector<int32_t> labels =
            initialize_labels(g, L.parts, P.hub_selection, stream_seed(seed, 2, depth));
        lap(1, t);
        const int sweeps = propagate_labels(g, labels, depth > 1);
        record(depth, g.n, g.edges(), sweeps);
        lap(2, t);
        label_outliers(g, labels, stream_seed(seed, 3, depth));
        const int64_t parts = remap_labels(labels);
        lap(3, t);
        const AggInput in = summed ? AggInput{summed->ptr.data(), summed->idx.data(), nullptr,
                                              summed->w.data()}
                                   : AggInput{g.ptr, g.idx, g.w, nullptr};
        Summed coarse_summed = aggregate(g.n, in, labels, parts);
        lap(4, t);
        Csr coarse_graph = transition(coarse_summed);
        lap(5, t);
        std::vector<float> coarse;
        {
            const Anchor ca = coarse_anchor(anchor, labels, parts);
            const std::optional<Level> next = next_level(L, parts, P);
            coarse = descend(coarse_graph.view(), &coarse_summed, ca, next, depth + 1);
        }
        coarse_summed = Summed();
        if (!coarse_graph.w.empty()) layout(coarse_graph, coarse, L, stream_seed(seed, 4, depth));
        t = Clock::now();
        UMAPHIP_REQUIRE(static_cast<int64_t>(coarse.size()) == parts * d, UMAPHIP_NUMERICAL,
                        "invalid recursive coarse layout output");
        for (float x : coarse)
            UMAPHIP_REQUIRE(std::isfinite(x), UMAPHIP_NUMERICAL,
                            "invalid recursive coarse layout output");
        coarse_graph = Csr();
        std::vector<float> result = prolongate(g, labels, coarse.data(), d);
        procrustes_align(result.data(), g.n, anchor);
        lap(6, t);
        return result;
    }
};

CsrView caller_view(int64_t n, const int64_t *offsets, const int32_t *indices,
                    const float *weights) {
    return {n, offsets, indices, weights};
}

Anchor caller_anchor(const float *values, const uint8_t *mask, int64_t n, int64_t d) {
    UMAPHIP_REQUIRE(values && mask && d > 0 && n >= 0, UMAPHIP_INVALID_INPUT,
                    "invalid recursive anchor");
    Anchor A;
    A.d = d;
    A.values.assign(values, values + n * d);
    A.mask.assign(mask, mask + n);
    for (int64_t r = 0; r < n; ++r)
        if (!A.mask[r])
            for (int64_t j = 0; j < d; ++j) A.values[r * d + j] = 0.0f;
    return A;
}

}  // namespace
}  // namespace umaphip

using namespace umaphip;

extern "C" {

umaphip_status umaphip_recursive_init(int64_t n, const int64_t *offsets, const int32_t *indices,
                                      const float *weights, const float *data, int64_t features,
                                      int64_t d, float a, float b,
                                      const umaphip_recursive_params *params, uint64_t seed,
                                      float *out, umaphip_recursive_stats *stats) {
    return guarded([&] {
        check_params(params);
        UMAPHIP_REQUIRE(n > 0 && d > 0 && features > 0 && data && out, UMAPHIP_INVALID_INPUT,
                        "invalid recursive initialization arguments");
        const CsrView g = caller_view(n, offsets, indices, weights);
        validate_graph(g);
        if (stats) *stats = umaphip_recursive_stats{};
        const auto t0 = Clock::now();
        Run run{*params, seed, d, stats};
        StdRng rng(stream_seed(seed, 1, 0));
        const Anchor anchor =
            anchor_reference(data, n, features, d, params->anchor_sample_size, rng);
        if (stats) stats->stage_ms[0] = ms_since(t0);
        const std::vector<float> result =
            run.descend(g, nullptr, anchor, first_level(n, a, b, *params), 1);
        std::memcpy(out, result.data(), result.size() * sizeof(float));
        if (stats) {
            stats->layout_ms = run.layout_ms;
            stats->host_ms = ms_since(t0) - run.layout_ms;
        }
    });
}

umaphip_status umaphip_recursive_level_params(int64_t n, const umaphip_recursive_level *previous,
                                       float a, float b, const umaphip_recursive_params *params,
                                       umaphip_recursive_level *out, int32_t *has_level) {
    return guarded([&] {
        check_params(params);
        UMAPHIP_REQUIRE(out && has_level && n >= 0, UMAPHIP_INVALID_INPUT,
                        "invalid recursive level arguments");
        std::optional<Level> L;
        if (previous) {
            Level prev{previous->parts, previous->epochs, previous->a, previous->b,
                       previous->good_initialization != 0, previous->curve_a, previous->curve_b};
            L = next_level(prev, n, *params);
        } else {
            L = first_level(n, a, b, *params);
        }
        *has_level = L ? 1 : 0;
        if (L) *out = {L->parts, L->epochs, L->a, L->b, L->good ? 1 : 0, L->curve_a, L->curve_b};
    });
}

umaphip_status umaphip_recursive_hubs(int64_t n, const int64_t *offsets, const int32_t *indices,
                                      const float *weights, int64_t parts, int32_t method,
                                      uint64_t seed, int32_t *labels) {
    return guarded([&] {
        const CsrView g = caller_view(n, offsets, indices, weights);
        check_shape(g);
        UMAPHIP_REQUIRE(labels && method >= 0 && method <= 2 && parts >= 0,
                        UMAPHIP_INVALID_INPUT, "invalid hub arguments");
        const auto v = initialize_labels(g, parts, method, seed);
        std::copy(v.begin(), v.end(), labels);
    });
}

umaphip_status umaphip_recursive_propagate(int64_t n, const int64_t *offsets,
                                           const int32_t *indices, const float *weights,
                                           int32_t share, int32_t *labels, int32_t *sweeps) {
    return guarded([&] {
        const CsrView g = caller_view(n, offsets, indices, weights);
        check_shape(g);
        UMAPHIP_REQUIRE(labels, UMAPHIP_INVALID_INPUT, "invalid propagation arguments");
        std::vector<int32_t> v(labels, labels + n);
        const int s = propagate_labels(g, v, share != 0);
        std::copy(v.begin(), v.end(), labels);
        if (sweeps) *sweeps = s;
    });
}

umaphip_status umaphip_recursive_outliers(int64_t n, const int64_t *offsets,
                                          const int32_t *indices, uint64_t seed,
                                          int32_t *labels) {
    return guarded([&] {
        const CsrView g{n, offsets, indices, nullptr};
        UMAPHIP_REQUIRE(labels && offsets && (offsets[n] == 0 || indices), UMAPHIP_INVALID_INPUT,
                        "invalid outlier arguments");
        for (int64_t e = 0; e < offsets[n]; ++e)
            UMAPHIP_REQUIRE(indices[e] >= 0 && indices[e] < n, UMAPHIP_INVALID_INPUT,
                            "invalid recursive CSR graph");
        std::vector<int32_t> v(labels, labels + n);
        label_outliers(g, v, seed);
        std::copy(v.begin(), v.end(), labels);
    });
}

umaphip_status umaphip_recursive_remap(int32_t *labels, int64_t n, int64_t *parts) {
    return guarded([&] {
        UMAPHIP_REQUIRE((labels || n == 0) && parts, UMAPHIP_INVALID_INPUT,
                        "invalid remap arguments");
        std::vector<int32_t> v(labels, labels + n);
        *parts = remap_labels(v);
        std::copy(v.begin(), v.end(), labels);
    });
}

umaphip_status umaphip_recursive_aggregate(int64_t n, const int64_t *offsets,
                                           const int32_t *indices, const double *weights,
                                           const int32_t *labels, int64_t parts,
                                           int64_t *out_offsets, int32_t *out_indices,
                                           double *out_weights) {
    return guarded([&] {
        UMAPHIP_REQUIRE(offsets && labels && out_offsets && parts >= 0, UMAPHIP_INVALID_INPUT,
                        "invalid aggregation arguments");
        for (int64_t r = 0; r < n; ++r)
            UMAPHIP_REQUIRE(labels[r] >= 0 && labels[r] < parts, UMAPHIP_INVALID_INPUT,
                            "labels must be compact part indices");
        for (int64_t e = 0; e < offsets[n]; ++e)
            UMAPHIP_REQUIRE(indices[e] >= 0 && indices[e] < n, UMAPHIP_INVALID_INPUT,
                            "invalid recursive CSR graph");
        const Summed s = aggregate(n, AggInput{offsets, indices, nullptr, weights},
                                   std::vector<int32_t>(labels, labels + n), parts);
        std::copy(s.ptr.begin(), s.ptr.end(), out_offsets);
        std::copy(s.idx.begin(), s.idx.end(), out_indices);
        std::copy(s.w.begin(), s.w.end(), out_weights);
    });
}

umaphip_status umaphip_recursive_transition(int64_t n, const int64_t *offsets,
                                            const int32_t *indices, const double *weights,
                                            int64_t *out_offsets, int32_t *out_indices,
                                            float *out_weights) {
    return guarded([&] {
        UMAPHIP_REQUIRE(offsets && out_offsets, UMAPHIP_INVALID_INPUT,
                        "invalid transition arguments");
        Summed s;
        s.n = n;
        s.ptr.assign(offsets, offsets + n + 1);
        s.idx.assign(indices, indices + offsets[n]);
        s.w.assign(weights, weights + offsets[n]);
        for (int64_t r = 0; r < n; ++r)
            for (int64_t e = s.ptr[r]; e < s.ptr[r + 1]; ++e)
                UMAPHIP_REQUIRE(s.idx[e] >= 0 && s.idx[e] < n &&
                                    (e == s.ptr[r] || s.idx[e - 1] < s.idx[e]),
                                UMAPHIP_INVALID_INPUT, "invalid summed graph");
        const Csr t = transition(s);
        std::copy(t.ptr.begin(), t.ptr.end(), out_offsets);
        std::copy(t.idx.begin(), t.idx.end(), out_indices);
        std::copy(t.w.begin(), t.w.end(), out_weights);
    });
}

umaphip_status umaphip_recursive_anchor(const float *data, int64_t n, int64_t features,
                                        int64_t d, int64_t sample_size, uint64_t seed,
                                        float *values, uint8_t *mask) {
    return guarded([&] {
        UMAPHIP_REQUIRE(values && mask, UMAPHIP_INVALID_INPUT, "invalid anchor outputs");
        StdRng rng(stream_seed(seed, 1, 0));
        const Anchor A = anchor_reference(data, n, features, d, sample_size, rng);
        std::copy(A.values.begin(), A.values.end(), values);
        std::copy(A.mask.begin(), A.mask.end(), mask);
    });
}

umaphip_status umaphip_recursive_base_case(int64_t n, const int64_t *offsets,
                                           const int32_t *indices, const float *weights,
                                           const float *anchor, const uint8_t *mask, int64_t d,
                                           int64_t depth, float *out) {
    return guarded([&] {
        const CsrView g = caller_view(n, offsets, indices, weights);
        check_shape(g);
        UMAPHIP_REQUIRE(out && n > 0, UMAPHIP_INVALID_INPUT, "invalid base case arguments");
        const auto v = base_case(g, caller_anchor(anchor, mask, n, d), depth);
        std::copy(v.begin(), v.end(), out);
    });
}

umaphip_status umaphip_recursive_prolongate(int64_t n, const int64_t *offsets,
                                            const int32_t *indices, const float *weights,
                                            const int32_t *labels, const float *coarse,
                                            int64_t parts, int64_t d, float *out) {
    return guarded([&] {
        const CsrView g = caller_view(n, offsets, indices, weights);
        check_shape(g);
        UMAPHIP_REQUIRE(labels && coarse && out && d > 0 && n > 0, UMAPHIP_INVALID_INPUT,
                        "invalid prolongation arguments");
        for (int64_t r = 0; r < n; ++r)
            UMAPHIP_REQUIRE(labels[r] >= 0 && labels[r] < parts, UMAPHIP_INVALID_INPUT,
                            "labels must be compact part indices");
        const auto v = prolongate(g, std::vector<int32_t>(labels, labels + n), coarse, d);
        std::copy(v.begin(), v.end(), out);
    });
}

umaphip_status umaphip_recursive_procrustes(float *layout, int64_t n, const float *anchor,
                                            const uint8_t *mask, int64_t d) {
    return guarded([&] {
        UMAPHIP_REQUIRE(layout && n > 0, UMAPHIP_INVALID_INPUT, "invalid Procrustes arguments");
        procrustes_align(layout, n, caller_anchor(anchor, mask, n, d));
    });
}

}  // extern "C"
