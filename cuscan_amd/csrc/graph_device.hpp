// Transform front-end on the device (graph_device.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace umaphip {

constexpr int kMaxGraphK = 64;

// Error flags of one request's graph (host message order: invalid ids, an
// empty graph, an initial row without positive memberships, a bad period).
constexpr int32_t kTGraphInvalidIds = 1;
constexpr int32_t kTGraphEmpty = 2;
constexpr int32_t kTGraphEmptyInitRow = 4;
constexpr int32_t kTGraphBadPeriod = 8;

struct TGraphState {
    float mean;            // f64 mean of the request's distances, as f32
    uint32_t wmax_bits;    // largest membership weight (bits of a nonnegative float)
    int64_t edges;         // kept edges
    int32_t errors;        // kTGraph* flags
    int32_t pad;
};

// One request: rows [row0, row0 + rows) of the batch's kNN arrays. Outputs
// are the request's own arrays (capacity rows * k edges).
struct TGraphSegment {
    int64_t row0, rows;
    uint32_t *offsets;                      // [rows + 1]
    uint32_t *indices;                      // [rows * k]
    double *periods, *negative_periods;     // [rows * k]
    double *next_positive, *next_negative;  // [rows * k]
    float *init;                            // [rows, d]
    TGraphState *state;
};

struct TGraphParams {
    uint32_t floor_conn, min_positive;
    float fraction, target;
    int32_t interpolate;
};

TGraphParams transform_graph_params(int64_t k, float local_connectivity);
// Segments in ascending row0 order; tmp_* hold total_rows * k (tail, weight)
// pairs and total_rows counts.
void launch_transform_graphs(const TGraphSegment *segs, int32_t n_segs, int64_t total_rows,
                             const int32_t *idx, const float *dist, int64_t k, int64_t n,
                             const TGraphParams &P, const float *embedding, int64_t d,
                             int32_t epochs, float rate, int32_t *tmp_tail, float *tmp_w,
                             int32_t *tmp_cnt, hipStream_t s);
const char *transform_graph_error(int32_t errors);
umaphip_status transform_graph_status(int32_t errors);

}  // namespace umaphip
