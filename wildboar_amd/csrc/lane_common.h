// lane_common.h -- helpers shared by the lane-per-pair kernels (lane.hip,
// train.hip): geometry passed to the kernels and the guarded y load.
#pragma once

#include "wbx_internal.h"

namespace wbx {

// One geometry for every pair of a launch: lengths, band offset and width,
// and the slot of the final cell (lx-1, ly-1).
struct LaneGeo {
    int lx, ly, off, B, kf;
};

// Plain DTW on a band of exactly B slots (odd B, 9 .. 111; lane_exact.hip).
bool lane_exact_supported(int B);
int launch_lane_exact(const PairSpec &ps, Series<float> xs, Series<float> ys, const LaneGeo &g,
                      float *sq, hipStream_t st);

// Four floats at a 4-byte aligned address (one 16-byte load).
typedef float f4a __attribute__((ext_vector_type(4), aligned(4)));

// y[j] for j in [0, ly), +inf outside (the cells there become +inf).
__device__ __forceinline__ float ldy(const float *__restrict__ Y, int j, int ly) {
    const int jc = j < 0 ? 0 : (j >= ly ? ly - 1 : j);
    const float v = Y[jc];
    return (unsigned)j < (unsigned)ly ? v : WBX_INF_F;
}

}  // namespace wbx
