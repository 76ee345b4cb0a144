// lane.hip -- launches of the lane-per-pair kernel (lane_kernel.h): bands
// padded to a multiple of 8 slots, every metric variant (DTW, WDTW, ADTW).
#include "lane_kernel.h"

namespace wbx {

// The y load for group i0 fetches window slots W-U .. W-1 of the *next* group:
// next window starts at i0 + U - off, its slot W-U+u is y[i0 + U - off + W-U+u].

template <int BP, int U>
static void launch_one(const PairSpec &ps, Series<float> xs, Series<float> ys, LaneGeo g,
                       const double *penalty, const float *weights, float *sq, hipStream_t st) {
    dim3 grid((unsigned)((ps.n_pairs + 63) / 64));
    if (penalty)
        hipLaunchKernelGGL((k_lane_dtw<BP, U, true, false>), grid, dim3(64), 0, st, ps, xs, ys, g,
                           (float)*penalty, nullptr, sq);
    else if (weights)
        hipLaunchKernelGGL((k_lane_dtw<BP, U, false, true>), grid, dim3(64), 0, st, ps, xs, ys, g,
                           0.0f, weights, sq);
    else
        hipLaunchKernelGGL((k_lane_dtw<BP, U, false, false>), grid, dim3(64), 0, st, ps, xs, ys, g,
                           0.0f, nullptr, sq);
}

constexpr int LANE_MAX_BP = 112;

bool lane_supported(int lx, int ly, double r, int weighted, int w_fixed) {
    (void)weighted;
    Band bd = make_band(lx, ly, r, w_fixed);
    return bd.B <= LANE_MAX_BP && lx >= 1 && ly >= 1;
}

int launch_lane_dtw(const PairSpec &ps, Series<float> xs, Series<float> ys, double r,
                    const float *weights, float *sq, hipStream_t st, int w_fixed,
                    const double *penalty) {
    if (ps.n_pairs <= 0) return WBX_OK;
    if (weights && penalty) return fail(WBX_EUNSUPPORTED, "lane kernel: weights with a penalty");
    if (xs.lens || ys.lens) return fail(WBX_EUNSUPPORTED, "lane kernel: needs one geometry");
    Band bd = make_band(xs.len, ys.len, r, w_fixed);
    if (bd.B > LANE_MAX_BP) return fail(WBX_EUNSUPPORTED, "lane kernel: band %d too wide", bd.B);
    LaneGeo g{bd.lx, bd.ly, bd.off, bd.B, bd.ly - bd.lx + bd.off};
    if (!weights && !penalty && lane_exact_supported(bd.B))
        return launch_lane_exact(ps, xs, ys, g, sq, st);
    const int BP = (bd.B + 7) / 8 * 8;
    switch (BP) {
        case 8: launch_one<8, 8>(ps, xs, ys, g, penalty, weights, sq, st); break;
        case 16: launch_one<16, 8>(ps, xs, ys, g, penalty, weights, sq, st); break;
        case 24: launch_one<24, 8>(ps, xs, ys, g, penalty, weights, sq, st); break;
        case 32: launch_one<32, 8>(ps, xs, ys, g, penalty, weights, sq, st); break;
        case 40: launch_one<40, 8>(ps, xs, ys, g, penalty, weights, sq, st); break;
        case 48: launch_one<48, 8>(ps, xs, ys, g, penalty, weights, sq, st); break;
        case 56: launch_one<56, 8>(ps, xs, ys, g, penalty, weights, sq, st); break;
        case 64: launch_one<64, 8>(ps, xs, ys, g, penalty, weights, sq, st); break;
        case 72: launch_one<72, 8>(ps, xs, ys, g, penalty, weights, sq, st); break;
        case 80: launch_one<80, 8>(ps, xs, ys, g, penalty, weights, sq, st); break;
        case 88: launch_one<88, 8>(ps, xs, ys, g, penalty, weights, sq, st); break;
        case 96: launch_one<96, 4>(ps, xs, ys, g, penalty, weights, sq, st); break;
        case 104: launch_one<104, 4>(ps, xs, ys, g, penalty, weights, sq, st); break;
        case 112: launch_one<112, 4>(ps, xs, ys, g, penalty, weights, sq, st); break;
        default: return fail(WBX_EUNSUPPORTED, "lane kernel: band %d", bd.B);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

}  // namespace wbx
