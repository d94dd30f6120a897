// Host random source reproducing the draws cuscan makes through rand's
// StdRng (rand 0.10.3: ChaCha12 with a 64-bit block counter and stream 0,
// seeded by rand_core's seed_from_u64 PCG32 expansion; chacha20 0.10.2).
//
// Call sites restated: ModernOrders shuffles
// (crates/umap/src/standard_umap/modern_orders.rs:359-397, 544-631),
// projection directions (modern_negatives.rs:283-297 via
// crates/umap-core/src/recursive_init/mod.rs:554-558) and adaptation samples
// (modern_negatives.rs:402-422).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>

namespace umaphip {

class StdRng {
public:
    explicit StdRng(uint64_t seed) { reseed(seed); }

    void reseed(uint64_t seed) {
        // PCG32 XSH-RR, one output word per 4 seed bytes.
        uint64_t s = seed;
        for (int w = 0; w < 8; ++w) {
            s = s * 6364136223846793005ULL + 11634580027462260723ULL;
            const uint32_t x = static_cast<uint32_t>(((s >> 18) ^ s) >> 27);
            const uint32_t r = static_cast<uint32_t>(s >> 59);
            key_[w] = (x >> r) | (x << ((32u - r) & 31u));
        }
        counter_ = 0;
        pos_ = kWords;
    }

    uint32_t next_u32() {
        if (pos_ == kWords) refill();
        return words_[pos_++];
    }

    uint64_t next_u64() {
        const uint64_t lo = next_u32();
        const uint64_t hi = next_u32();
        return lo | (hi << 32);
    }

    // Uniform double in [0, 1) from the top 53 bits.
    double next_f64() { return static_cast<double>(next_u64() >> 11) * 0x1.0p-53; }

    // Uniform integer in [0, bound) (Canon's widening multiply, one
    // follow-up word when the low product half could carry).
    uint32_t below(uint32_t bound) {
        const uint64_t wide = static_cast<uint64_t>(next_u32()) * bound;
        uint32_t hi = static_cast<uint32_t>(wide >> 32);
        const uint32_t lo = static_cast<uint32_t>(wide);
        if (lo > 0u - bound) {
            const uint64_t wide2 = static_cast<uint64_t>(next_u32()) * bound;
            if (static_cast<uint64_t>(lo) + (wide2 >> 32) > 0xFFFFFFFFull) hi += 1;
        }
        return hi;
    }

    // Uniform f32 in [low, high): 23 mantissa bits in [1, 2), minus one,
    // times the span, plus low; the span shrinks by one ulp in the
    // measure-zero case that rounding reaches high.
    float range_f32(float low, float high) {
        float span = high - low;
        for (;;) {
            const uint32_t bits = (next_u32() >> 9) | 0x3F800000u;
            float unit;
            std::memcpy(&unit, &bits, 4);
            const float scaled = (unit - 1.0f) * span;
            const float value = scaled + low;
            if (value < high) return value;
            uint32_t sb;
            std::memcpy(&sb, &span, 4);
            sb -= 1;
            std::memcpy(&span, &sb, 4);
        }
    }

    // Box-Muller pair exactly as cuscan's gaussian_pair.
    void gaussian_pair(double &x, double &y) {
        const double r = std::sqrt(-2.0 * std::log(1.0 - next_f64()));
        const double t = 6.283185307179586 * next_f64();
        x = r * std::cos(t);
        y = r * std::sin(t);
    }

    // SliceRandom::shuffle: Durstenfeld pass i = 0..n, swap(i, idx) with idx
    // drawn from an IncreasingUniform chooser that packs several bounded
    // draws into one u32. `permute == false` consumes exactly the same
    // words without touching `a` (used where only the generator state
    // matters).
    void shuffle(int32_t *a, size_t n, bool permute = true) {
        if (n <= 1) return;
        if (!permute) {
            // The words drawn depend only on the chunk boundaries: element 0
            // draws nothing, and a chunk starting at element i (span i + 1)
            // draws below(product) once and covers `factors` elements.
            for (size_t i = 1; i < n;) {
                uint32_t product, factors;
                chunk_bound(static_cast<uint32_t>(i + 1), product, factors);
                (void)below(product);
                i += factors;
            }
            return;
        }
        uint32_t seen = 0, packed = 0, left = 1;
        for (size_t i = 0; i < n; ++i) {
            const uint32_t span = seen + 1;
            if (left == 0) {
                uint32_t product, factors;
                chunk_bound(span, product, factors);
                packed = below(product);
                left = factors;
            }
            --left;
            uint32_t pick;
            if (left == 0) {
                pick = packed;
            } else {
                pick = packed % span;
                packed /= span;
            }
            seen = span;
            if (permute) {
                const int32_t t = a[i];
                a[i] = a[pick];
                a[pick] = t;
            }
        }
    }

private:
    static constexpr unsigned kWords = 64;  // four ChaCha blocks per refill

    static void chunk_bound(uint32_t m, uint32_t &product, uint32_t &factors) {
        uint64_t p = m, next = static_cast<uint64_t>(m) + 1;
        while (p * next <= 0xFFFFFFFFull) {
            p *= next;
            ++next;
        }
        product = static_cast<uint32_t>(p);
        factors = static_cast<uint32_t>(next - m);
    }

    // The four blocks of a refill are computed side by side: word w of the
    // four blocks is one 4-lane vector (GCC/Clang vector extension; SSE2 on
    // x86-64), so every step of a quarter round is one vector operation. The
    // words are those of four separate block computations.
    typedef uint32_t v4u __attribute__((vector_size(16)));

    static inline v4u rotv(v4u v, int c) { return (v << c) | (v >> (32 - c)); }

    static inline void quarter4(v4u *x, int a, int b, int c, int d) {
        x[a] += x[b]; x[d] = rotv(x[d] ^ x[a], 16);
        x[c] += x[d]; x[b] = rotv(x[b] ^ x[c], 12);
        x[a] += x[b]; x[d] = rotv(x[d] ^ x[a], 8);
        x[c] += x[d]; x[b] = rotv(x[b] ^ x[c], 7);
    }

    void refill() {
        v4u in[16], x[16];
        const uint32_t sigma[4] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        for (int w = 0; w < 4; ++w) in[w] = v4u{sigma[w], sigma[w], sigma[w], sigma[w]};
        for (int w = 0; w < 8; ++w) in[4 + w] = v4u{key_[w], key_[w], key_[w], key_[w]};
        for (int l = 0; l < 4; ++l) {
            const uint64_t c = counter_ + static_cast<uint64_t>(l);
            in[12][l] = static_cast<uint32_t>(c);
            in[13][l] = static_cast<uint32_t>(c >> 32);
        }
        in[14] = v4u{0u, 0u, 0u, 0u};
        in[15] = v4u{0u, 0u, 0u, 0u};
        for (int w = 0; w < 16; ++w) x[w] = in[w];
        for (int round = 0; round < 6; ++round) {  // 12 rounds
            quarter4(x, 0, 4, 8, 12); quarter4(x, 1, 5, 9, 13);
            quarter4(x, 2, 6, 10, 14); quarter4(x, 3, 7, 11, 15);
            quarter4(x, 0, 5, 10, 15); quarter4(x, 1, 6, 11, 12);
            quarter4(x, 2, 7, 8, 13); quarter4(x, 3, 4, 9, 14);
        }
        for (int w = 0; w < 16; ++w) {
            const v4u o = x[w] + in[w];
            for (int l = 0; l < 4; ++l) words_[l * 16 + w] = o[l];
        }
        counter_ += 4;
        pos_ = 0;
    }

    uint32_t key_[8];
    uint64_t counter_ = 0;
    uint32_t words_[kWords];
    unsigned pos_ = kWords;
};

}  // namespace umaphip
