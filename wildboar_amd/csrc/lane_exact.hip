// lane_exact.hip -- the lane-per-pair DTW kernel (lane_kernel.h) compiled for
// every odd band width B = 9 .. 111 exactly.  Two series of equal length have
// B = 2w - 1 (|lx - ly| + 2w - 1 in general), so pairwise and paired DTW on
// a fixed-width array and subsequence windows run on these: the padded
// kernels would compute up to 7 slots per row outside the band and reset one
// of them every row.
#include "lane_kernel.h"

namespace wbx {

namespace {

template <int B>
void launch_exact(const PairSpec &ps, Series<float> xs, Series<float> ys, const LaneGeo &g,
                  float *sq, hipStream_t st) {
    // rows per window shift as the padded kernels use for the same width
    constexpr int U = (B + 7) / 8 * 8 <= 88 ? 8 : 4;
    dim3 grid((unsigned)((ps.n_pairs + 63) / 64));
    hipLaunchKernelGGL((k_lane_dtw_exact<B, U>), grid, dim3(64), 0, st, ps, xs, ys, g, sq);
}

}  // namespace

bool lane_exact_supported(int B) { return B >= 9 && B <= 111 && (B & 1); }

int launch_lane_exact(const PairSpec &ps, Series<float> xs, Series<float> ys, const LaneGeo &g,
                      float *sq, hipStream_t st) {
    switch (g.B) {
#define WBX_EXACT(b) \
    case b: launch_exact<b>(ps, xs, ys, g, sq, st); break;
        WBX_EXACT(9) WBX_EXACT(11) WBX_EXACT(13) WBX_EXACT(15) WBX_EXACT(17) WBX_EXACT(19)
        WBX_EXACT(21) WBX_EXACT(23) WBX_EXACT(25) WBX_EXACT(27) WBX_EXACT(29) WBX_EXACT(31)
        WBX_EXACT(33) WBX_EXACT(35) WBX_EXACT(37) WBX_EXACT(39) WBX_EXACT(41) WBX_EXACT(43)
        WBX_EXACT(45) WBX_EXACT(47) WBX_EXACT(49) WBX_EXACT(51) WBX_EXACT(53) WBX_EXACT(55)
        WBX_EXACT(57) WBX_EXACT(59) WBX_EXACT(61) WBX_EXACT(63) WBX_EXACT(65) WBX_EXACT(67)
        WBX_EXACT(69) WBX_EXACT(71) WBX_EXACT(73) WBX_EXACT(75) WBX_EXACT(77) WBX_EXACT(79)
        WBX_EXACT(81) WBX_EXACT(83) WBX_EXACT(85) WBX_EXACT(87) WBX_EXACT(89) WBX_EXACT(91)
        WBX_EXACT(93) WBX_EXACT(95) WBX_EXACT(97) WBX_EXACT(99) WBX_EXACT(101) WBX_EXACT(103)
        WBX_EXACT(105) WBX_EXACT(107) WBX_EXACT(109) WBX_EXACT(111)
#undef WBX_EXACT
        default: return fail(WBX_EUNSUPPORTED, "lane kernel: no exact instantiation for band %d", g.B);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

}  // namespace wbx
