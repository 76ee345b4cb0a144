// lane64_cost.h -- the float64 register-band DTW / WDTW / ADTW cost of one
// pair or window per lane (slot k of row i <-> column j = i + k - off; up =
// slot k+1 of the previous row, left = slot k-1 of this row, diag = slot k of
// the previous row), with the reference's cell arithmetic
// (distance/_elastic.pyx:925-928, contraction off):
//     v = x_i - y_j ;  c = min(min(up, left), diag) + (v * v) * w
// bit for bit.  Used by the lane64 kernels (pairs, window profiles,
// subsequence minimum) and by the float64 finish of the FP32 subsequence
// scan's candidates.
#pragma once
#pragma clang fp contract(off)

#include "wbx_internal.h"

namespace wbx {

namespace {

__device__ __forceinline__ double ldy64(const double *__restrict__ Y, int j, int ly) {
    const int jc = j < 0 ? 0 : (j >= ly ? ly - 1 : j);
    const double v = Y[jc];
    return (unsigned)j < (unsigned)ly ? v : WBX_INF_D;
}

// The reference's min, b < a ? b : a.  F: as one v_min_f64, equal here --
// the operands are never NaN (finite inputs; +inf only meets finite values
// or +inf, the penalty is > -inf) and never -0 (costs are sums of squares,
// +0 + -0 is +0), where the two could differ.  The wide rows keep the
// compare-and-select form: with v_min_f64 the scheduler keeps more values
// live and spills them to scratch.
template <bool F>
__device__ __forceinline__ double dmin(double a, double b) {
    if constexpr (F)
        return __builtin_fmin(a, b);
    else
        return b < a ? b : a;
}

template <int BP, bool AD>
constexpr bool fast_min() {
    return AD ? BP <= 56 : BP <= 80;
}

// One row of BP cells.  wk[k] is the weight of slot k (unused when !WT).
// AD: the amercing recursion, min(min(up + p, left + p), diag) + v * v
// (adtw_distance, distance/_elastic.pyx:943-1008); row 0 passes p = 0, which
// leaves the prefix sum exact (x + 0 == x for the costs, all >= 0 or +inf).
template <int BP, bool WT, bool AD>
__device__ __forceinline__ void row64(double (&row)[BP], const double *yw, double xi,
                                      const double (&pen)[8], const double *wk, double p) {
    double left = WBX_INF_D;
#pragma unroll
    for (int k = 0; k < BP; ++k) {
        const double up = (k + 1 < BP) ? row[k + 1] : WBX_INF_D;
        double m;
        if constexpr (AD)
            m = dmin<fast_min<BP, AD>()>(dmin<fast_min<BP, AD>()>(up + p, left + p), row[k]);
        else
            m = dmin<fast_min<BP, AD>()>(dmin<fast_min<BP, AD>()>(up, left), row[k]);
        const double v = xi - yw[k];
        double term = v * v;
        if constexpr (WT) term = term * wk[k];
        const double c = m + term;
        row[k] = c;
        left = c;
    }
#pragma unroll
    for (int q = 1; q < 8; ++q) row[BP - 8 + q] = fmax(row[BP - 8 + q], pen[q]);
}

// Squared DTW cost of X[0, lx) against Y[0, ly) with band offset `off`,
// width B <= BP, final slot kf.  weights: wn >= max(lx, ly) values, all > 0.
// Whether a kernel's register class loads the next group's x and window
// entries during the current group: the classes where the extra registers
// cost no wave per SIMD and no spills (the others issue each group's loads at
// its start).  kind 0: pairs, 1: window profiles, 2: subsequence minimum.
template <int KIND, int BP, bool WT, bool AD>
constexpr bool lane64_prefetch() {
    if constexpr (KIND == 0)
        return WT ? (BP == 16 || BP == 32 || BP == 48)
                  : AD ? (BP >= 24 && BP <= 40) : (BP == 16 || (BP >= 32 && BP <= 56) || BP == 72);
    else if constexpr (KIND == 1)
        return BP <= (WT ? 56 : 48);
    else
        return WT || AD ? BP <= 40 : (BP == 8 || (BP >= 32 && BP <= 48));
}

template <int BP, int U, bool WT, bool AD = false, bool PF = false>
__device__ __forceinline__ double lane64_cost(const double *__restrict__ X, int lx,
                                              const double *__restrict__ Y, int ly, int off,
                                              int B, int kf,
                                              const double *__restrict__ weights, int wn,
                                              double p = 0.0) {
    constexpr int W = BP + U - 1;
    double row[BP];
#pragma unroll
    for (int k = 0; k < BP; ++k) row[k] = (k == off) ? 0.0 : WBX_INF_D;
    double pen[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) pen[q] = (BP - 8 + q == B) ? WBX_INF_D : -WBX_INF_D;
    double wk[WT ? BP : 1];
    double w0[WT ? BP : 1];
    if constexpr (WT) {
#pragma unroll
        for (int k = 0; k < BP; ++k) {
            int a = k - off;
            // row 0 weighs column j > 0 with w[j - 1], column 0 with w[0]
            int a0 = a > 0 ? a - 1 : 0;
            a = a < 0 ? -a : a;
            a = a < wn ? a : wn - 1;
            a0 = a0 < wn ? a0 : wn - 1;
            wk[k] = weights[a];
            w0[k] = weights[a0];
        }
    }
    double yw[W];
#pragma unroll
    for (int m = 0; m < W; ++m) yw[m] = ldy64(Y, m - off, ly);

    // row 0 on its own (its weights differ), then single rows until the rest
    // is a whole number of U-groups
    row64<BP, WT, AD>(row, yw, X[0], pen, w0, 0.0);
#pragma unroll
    for (int m = 0; m < W - 1; ++m) yw[m] = yw[m + 1];
    yw[W - 1] = ldy64(Y, 1 - off + W - 1, ly);
    const int rem = (lx - 1) % U;
    for (int i = 1; i <= rem; ++i) {
        row64<BP, WT, AD>(row, yw, X[i], pen, wk, p);
#pragma unroll
        for (int m = 0; m < W - 1; ++m) yw[m] = yw[m + 1];
        yw[W - 1] = ldy64(Y, i + 1 - off + W - 1, ly);
    }
    if constexpr (!PF) {
        for (int i0 = 1 + rem; i0 < lx; i0 += U) {
            double xv[U], yn[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                xv[u] = X[i0 + u];
                yn[u] = ldy64(Y, i0 + U - off + W - U + u, ly);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) row64<BP, WT, AD>(row, yw + u, xv[u], pen, wk, p);
#pragma unroll
            for (int m = 0; m < W - U; ++m) yw[m] = yw[m + U];
#pragma unroll
            for (int u = 0; u < U; ++u) yw[W - U + u] = yn[u];
        }
    } else {
        // x of the current group of U rows; the loop loads the next group's
        // with the group's new window entries (columns jn .. jn + U - 1,
        // jn > 0), so no row waits on a load issued in the same group.  The
        // lanes share lx, ly and the band: when both lie inside (every group
        // but the last one or two) they are plain 16-byte loads, else x is
        // clamped into its row (never used past it) and y is +inf outside.
        static_assert(U % 2 == 0, "groups of whole 16-byte loads");
        typedef double d2a __attribute__((ext_vector_type(2), aligned(8)));
        double xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = X[min(1 + rem + u, lx - 1)];
        for (int i0 = 1 + rem; i0 < lx; i0 += U) {
            double xn[U], yn[U];
            const int jn = i0 - off + W;
            if (i0 + 2 * U <= lx && jn + U <= ly) {
#pragma unroll
                for (int u = 0; u < U; u += 2) {
                    const d2a a = *(const d2a *)(X + i0 + U + u);
                    const d2a b = *(const d2a *)(Y + jn + u);
                    xn[u] = a[0];
                    xn[u + 1] = a[1];
                    yn[u] = b[0];
                    yn[u + 1] = b[1];
                }
            } else {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    xn[u] = X[min(i0 + U + u, lx - 1)];
                    yn[u] = ldy64(Y, jn + u, ly);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) row64<BP, WT, AD>(row, yw + u, xv[u], pen, wk, p);
#pragma unroll
            for (int m = 0; m < W - U; ++m) yw[m] = yw[m + U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                yw[W - U + u] = yn[u];
                xv[u] = xn[u];
            }
        }
    }
    double res = WBX_INF_D;
#pragma unroll
    for (int k = 0; k < BP; ++k)
        if (k == kf) res = row[k];
    return res;
}

}  // namespace

}  // namespace wbx
