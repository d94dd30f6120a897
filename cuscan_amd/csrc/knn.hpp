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

private:
    int64_t n_, d_, dp_ = 0, reserved_q_ = 0;
    int ks_ = 0;
    DeviceArray<float> ref_, rnorm_, cand_d_, top_d_;
    DeviceArray<float> sample_d_;     // [q, 8192] sample-bound distances of a batch
    DeviceArray<uint16_t> ref_bf_, qbf_;
    DeviceArray<int32_t> cand_i_, cand_n_, top_i_, top_n_;
    DeviceArray<uint32_t> tau_key_;   // per query: best published slice threshold (float key)
};

}  // namespace umaphip
