// Resident reference set for the exact kNN (BF16 prefilter + FP32 rerank).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace umaphip {

class KnnIndex {
public:
    // Copies the f32 references (host or device pointer) and builds the
    // BF16 copy (rows padded to a multiple of 32 features) and its norms.
    KnnIndex(const float *ref_host, const float *ref_device, int64_t n, int64_t d, int64_t max_q,
             hipStream_t s);
    // Grows the query workspace to hold batches of up to q rows.
    void reserve(int64_t q);
    // k <= 64 exact neighbours of q_dev [nq, d] (device f32).
    // With cand_ids set, writes the merged prefilter lists instead ([nq, 64]
    // ids best first, counts [nq]) and skips the rerank.
    void query(const float *q_dev, int64_t nq, int k, bool exclude_self, int64_t query_offset,
               int32_t *idx_out, float *dist_out, hipStream_t s,
               const float *q_f32_for_rerank = nullptr, int32_t *cand_ids = nullptr,
               int32_t *cand_counts = nullptr);
    int64_t rows() const { return n_; }
    int64_t dims() const { return d_; }
    const float *reference() const { return ref_.ptr; }
    int slices_for(int64_t q) const;
    // Whether a batch of q rows seeds its thresholds from the sample bound.
    bool uses_sample(int64_t q) const;
    int sample_size(int64_t q) const;
    // Queries of the last query() that the certificate left open and the
    // exact FP32 rescan answered (synchronises s).
    int64_t rescans(hipStream_t s) const;
    // eps: every reference row's BF16 rounding is at most eps times its
    // BF16 norm (0 when the shifted references are exact in BF16).
    float rounding_ratio(hipStream_t s) const;
    // Tests: rescan every query exactly (the certificate is not consulted).
    void set_force_rescan(bool on) { force_rescan_ = on; }

private:
    int64_t n_, d_, dp_ = 0, reserved_q_ = 0;
    int ks_ = 0, ex_slices_ = 1;
    bool force_rescan_ = false;
    DeviceArray<float> shift_;        // [d] per-feature shift subtracted before BF16
    DeviceArray<uint32_t> eps_key_;   // reference rounding ratio (float bits)
    DeviceArray<float> qnorm_, qerr_, kth_;   // per query: |x~|^2, rounding, k-th exact d^2
    DeviceArray<int32_t> fail_list_, fail_count_;
    DeviceArray<float> ex_d_;         // exact rescan: per (query slot, slice) best 64
    DeviceArray<int32_t> ex_i_, ex_n_;
    DeviceArray<float> ref_, rnorm_, cand_d_, top_d_;
    DeviceArray<float> sample_d_;     // [q, 8192] sample-bound distances of a batch
    DeviceArray<uint16_t> ref_bf_, qbf_;
    DeviceArray<int32_t> cand_i_, cand_n_, top_i_, top_n_;
    DeviceArray<uint32_t> tau_key_;   // per query: best published slice threshold (float key)
};

}  // namespace umaphip
