// lane_kernel.h -- FP32 DTW with one pair per lane and the band row in
// registers: the kernel template shared by the padded-band instantiations
// (lane.hip) and the exact-band ones (lane_exact.hip).
//
// For pairs that share one geometry (lx, ly, w) -- every pair of a
// pairwise/paired call on a fixed-width array, every window of one shapelet --
// each of the 64 lanes of a wave runs the whole DP of its own pair.  The row
// is kept in band coordinates (slot k <-> column j = i + k - off), where the
// three predecessors of slot k are slot k+1 of the previous row (up), slot
// k-1 of this row (left) and slot k of the previous row (diag).  Those are
// static register indices, so a fully unrolled row is BP cells of
//     d = x_i - y_j ; m = min3(up, left, diag) ; c = fma(d, d, m)
// i.e. three VALU instructions per cell and no LDS or cross-lane traffic.
//
// The y values of a row are a window y[i-off .. i-off+BP) that slides by one
// per row.  The kernel keeps a window of BP+U-1 registers, runs U rows against
// it (row u of the group reads yw[k+u]) and then shifts it by U: the shift
// costs (BP+U-1)/U moves per U*BP cells.  Columns outside [0, ly) read +inf,
// which makes their cells +inf, and the row state before row 0 is +inf with a
// 0 in slot off, which turns the first row into the reference's prefix sum.
// Padded instantiations round the band up to BP, a multiple of 8: slots >= B
// compute throw-away values and slot B is reset to +inf after every row so
// slot B-1 never sees them.  Exact instantiations (BP == B, odd B, the band
// of every pair of equal lengths) compute no padding slot and need no reset.
//
// WDTW (WeightedDtwMetric, distance/_elastic.pyx:3322-3341) multiplies each
// cell's squared difference by w[|i - j|]; in band coordinates |i - j| =
// |k - off| is fixed per slot, so every lane of a launch uses the same weight
// per slot: c = fma(d * d, wslot[k], m), four VALU instructions per cell.
// The per-slot weights are staged in LDS once per block and re-read every row
// (broadcast reads, no VGPRs held across rows).  Row 0 uses the reference's
// prefix-sum weights w[j - 1] (w[0] for j = 0, :893-906), a second LDS row.
#pragma once

#include "lane_common.h"

namespace wbx {

// pen[q] is +inf for slot BP-8+q == B and -inf otherwise (uniform values), so
// max(row, pen) resets slot B without a branch.  EX: the band is exactly BP
// slots (B == BP), the last slot's up input is the constant +inf and no slot
// needs a reset.
// AD: the amercing recursion (adtw_distance, distance/_elastic.pyx:943-1008),
// m = min3(up + p, left + p, diag); row 0 runs with p = 0 (its prefix sum
// carries no penalty).
template <int BP, bool AD = false, bool WT = false, bool EX = false>
__device__ __forceinline__ void lane_row(float (&row)[BP], const float *yw, float xi,
                                         const float (&pen)[8], float p = 0.0f,
                                         const float *wslot = nullptr) {
    float left = WBX_INF_F;
#pragma unroll
    for (int k = 0; k < BP; ++k) {
        const float up = (k + 1 < BP) ? row[k + 1] : WBX_INF_F;
        float m;
        if constexpr (AD)
            m = fminf(fminf(up + p, left + p), row[k]);
        else
            m = fminf(fminf(up, left), row[k]);
        const float d = xi - yw[k];
        float c;
        if constexpr (WT)
            c = __builtin_fmaf(d * d, wslot[k], m);
        else
            c = __builtin_fmaf(d, d, m);
        row[k] = c;
        left = c;
    }
    if constexpr (!EX) {
#pragma unroll
        for (int q = 1; q < 8; ++q) row[BP - 8 + q] = fmaxf(row[BP - 8 + q], pen[q]);
    }
}

template <int BP, int U, bool AD = false, bool WT = false, bool EX = false>
__device__ __forceinline__ void lane_dtw_body(PairSpec ps, Series<float> xs,
                                                 Series<float> ys, LaneGeo g, float pen_ad,
                                                 const float *__restrict__ weights,
                                                 float *__restrict__ sq) {
    constexpr int W = BP + U - 1;
    // WT: lw[0][k] row-0 weight of slot k, lw[1][k] = w[|k - off|]; slots
    // outside the band get 1 (their cells are +inf or thrown away)
    __shared__ float lw[WT ? 2 : 1][WT ? BP : 1];
    if constexpr (WT) {
        for (int k = threadIdx.x; k < BP; k += 64) {
            const int j = k - g.off;  // column of slot k in row 0
            lw[0][k] = (j >= 0 && k < g.B) ? weights[j > 0 ? j - 1 : 0] : 1.0f;
            lw[1][k] = k < g.B ? weights[j < 0 ? -j : j] : 1.0f;
        }
        __syncthreads();
    }
    const int64_t p0 = xcd_block(blockIdx.x, gridDim.x) * 64 + threadIdx.x;
    const int64_t p = p0 < ps.n_pairs ? p0 : ps.n_pairs - 1;
    int64_t a, b;
    decode_pair(ps, p, a, b);
    const float *__restrict__ X = xs.row(a);
    const float *__restrict__ Y = ys.row(b);
    const int lx = g.lx, ly = g.ly, off = g.off, B = g.B;

    float row[BP];
#pragma unroll
    for (int k = 0; k < BP; ++k) row[k] = (k == off) ? 0.0f : WBX_INF_F;
    float pen[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) pen[q] = (BP - 8 + q == B) ? WBX_INF_F : -WBX_INF_F;

    float yw[W];
#pragma unroll
    for (int m = 0; m < W; ++m) yw[m] = ldy(Y, m - off, ly);

    int i_begin = 0;
    if constexpr (WT) {
        lane_row<BP, false, true, EX>(row, yw, X[0], pen, 0.0f, lw[0]);
#pragma unroll
        for (int m = 0; m < W - 1; ++m) yw[m] = yw[m + 1];
        yw[W - 1] = ldy(Y, 1 - off + W - 1, ly);
        i_begin = 1;
    }
    if constexpr (AD) {
        lane_row<BP, true, false, EX>(row, yw, X[0], pen, 0.0f);
#pragma unroll
        for (int m = 0; m < W - 1; ++m) yw[m] = yw[m + 1];
        yw[W - 1] = ldy(Y, 1 - off + W - 1, ly);
        i_begin = 1;
    }
    // Leading rows one at a time so the rest is a whole number of U-groups.
    const int rem = (lx - i_begin) % U;
    for (int i = i_begin; i < i_begin + rem; ++i) {
        lane_row<BP, AD, WT, EX>(row, yw, X[i], pen, pen_ad, WT ? lw[WT ? 1 : 0] : nullptr);
#pragma unroll
        for (int m = 0; m < W - 1; ++m) yw[m] = yw[m + 1];
        yw[W - 1] = ldy(Y, i + 1 - off + W - 1, ly);
    }
    if constexpr (BP >= 112 || (WT && BP > 40) || (AD && BP >= 80)) {
        // the widest class, weighted bands past 40 slots and amerced ones
        // from 80 (their weights and penalties take registers too) have none
        // left for the next group's x: each group's loads are issued at its
        // start
        for (int i0 = i_begin + rem; i0 < lx; i0 += U) {
            float xv[U], yn[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                xv[u] = X[i0 + u];
                yn[u] = ldy(Y, i0 + U - off + W - U + u, ly);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if constexpr (WT) asm volatile("" ::: "memory");
                lane_row<BP, AD, WT, EX>(row, yw + u, xv[u], pen, pen_ad,
                                         WT ? lw[WT ? 1 : 0] : nullptr);
            }
#pragma unroll
            for (int m = 0; m < W - U; ++m) yw[m] = yw[m + U];
#pragma unroll
            for (int u = 0; u < U; ++u) yw[W - U + u] = yn[u];
        }
    } else {
        // x of the current group of U rows; the loop loads the next group's
        // with the group's new window entries (columns jn .. jn + U - 1,
        // jn > 0), so nothing a row reads waits on a load issued in the same
        // group.  The lanes share the geometry: when both lie inside their
        // rows (every group but the last one or two) they are plain 16-byte
        // loads, else x is clamped into the row (never used past it) and y is
        // +inf outside it.
        static_assert(U % 4 == 0, "groups of whole 16-byte loads");
        float xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = X[min(i_begin + rem + u, lx - 1)];
        for (int i0 = i_begin + rem; i0 < lx; i0 += U) {
            float xn[U], yn[U];
            const int jn = i0 - off + W;
            if (i0 + 2 * U <= lx && jn + U <= ly) {
#pragma unroll
                for (int u = 0; u < U; u += 4) {
                    const f4a a = *(const f4a *)(X + i0 + U + u);
                    const f4a b = *(const f4a *)(Y + jn + u);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        xn[u + q] = a[q];
                        yn[u + q] = b[q];
                    }
                }
            } else {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    xn[u] = X[min(i0 + U + u, lx - 1)];
                    yn[u] = ldy(Y, jn + u, ly);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                // the weights are re-read from LDS every row rather than held
                // in registers across the loop (the band row and y window
                // fill them)
                if constexpr (WT) asm volatile("" ::: "memory");
                lane_row<BP, AD, WT, EX>(row, yw + u, xv[u], pen, pen_ad,
                                         WT ? lw[WT ? 1 : 0] : nullptr);
            }
#pragma unroll
            for (int m = 0; m < W - U; ++m) yw[m] = yw[m + U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                yw[W - U + u] = yn[u];
                xv[u] = xn[u];
            }
        }
    }

    float res = WBX_INF_F;
#pragma unroll
    for (int k = 0; k < BP; ++k)
        if (k == g.kf) res = row[k];
    if (p0 < ps.n_pairs) sq[p0] = res;
}

// Plain DTW on padded bands is held to two waves per SIMD like the exact
// kernels (at 104 and 112 slots the allocator would take one wave's worth);
// the weighted and amerced variants need more registers than two waves have.
template <int BP, int U, bool AD = false, bool WT = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(AD || WT ? 1 : 2))) void
k_lane_dtw(PairSpec ps, Series<float> xs, Series<float> ys, LaneGeo g, float pen_ad,
           const float *__restrict__ weights, float *__restrict__ sq) {
    lane_dtw_body<BP, U, AD, WT, false>(ps, xs, ys, g, pen_ad, weights, sq);
}

// Exact bands, plain DTW.  Two waves per SIMD (at most 256 registers) is what
// keeps the VALU busy -- one wave issues about half as often -- so the
// allocator is held to it; the band row and y window fit (2B + 3 registers).
template <int B, int U>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_lane_dtw_exact(
    PairSpec ps, Series<float> xs, Series<float> ys, LaneGeo g, float *__restrict__ sq) {
    lane_dtw_body<B, U, false, false, true>(ps, xs, ys, g, 0.0f, nullptr, sq);
}

}  // namespace wbx
