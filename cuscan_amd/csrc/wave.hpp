// Device-side wave primitives shared by the kernels (64-lane waves).
#pragma once

#include <hip/hip_runtime.h>

namespace umaphip {

// Inclusive prefix sum over the wave's 64 lanes through DPP (row_shr:1..8
// within 16-lane rows, then row_bcast:15 and row_bcast:31 across them): a
// few cycles per step instead of the LDS round trips of lane permutes.
__device__ inline int wave_inclusive_sum(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);   // row_bcast:15 into rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);   // row_bcast:31 into rows 2, 3
    return v;
}

}  // namespace umaphip
