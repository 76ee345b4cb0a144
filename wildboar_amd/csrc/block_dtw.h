// block_dtw.h -- the reference's banded DTW recursion as a block-level
// anti-diagonal sweep (used by the exact pairwise and subsequence kernels).
// Arithmetic order per cell is the reference's (distance/_elastic.pyx:925-928):
//     cost = min(min(up, left), diag) + (v * v) * w,   v = x_i - y_j
// Files including this header compile with contraction off.
#pragma once

#include "wbx_internal.h"

namespace wbx {

__device__ __forceinline__ int floordiv2(int a) { return a >= 0 ? a / 2 : -((1 - a) / 2); }

template <typename T>
__device__ __forceinline__ T inf_of();
template <>
__device__ __forceinline__ float inf_of<float>() { return WBX_INF_F; }
template <>
__device__ __forceinline__ double inf_of<double>() { return WBX_INF_D; }

// Cython's min() on C doubles: (b < a) ? b : a.
template <typename T>
__device__ __forceinline__ T cmin(T a, T b) { return b < a ? b : a; }

// ------------------------------------------------------------- plain DTW
// Diagonal d = i + j holds cells i in [ilo(d), ihi(d)] (band and matrix
// clipped).  Each diagonal also writes +inf just outside that range, so every
// neighbour read of the next two diagonals sees either a band cell or +inf.
// Returns the squared cost of cell (lx-1, ly-1) in every lane.
// AD: the amercing recursion of adtw_distance (distance/_elastic.pyx:943-1008):
//     cost = min(min(up + p, left + p), diag) + v * v   (rows i >= 1)
// with row 0 the plain prefix sum (no penalty, no weights).
template <typename T, bool AD = false>
__device__ T block_dtw(const T *__restrict__ X, int lx, const T *__restrict__ Y, int ly, int w,
                       const T *__restrict__ weights, T *d0, T *d1, T *d2, T pen = (T)0) {
    const int min_len = lx > ly ? lx - ly : 0;
    const int max_len = (ly > lx ? ly - lx : 0) + w;
    const T INF = inf_of<T>();
    const int tid = threadIdx.x;
    for (int d = 0; d <= lx + ly - 2; ++d) {
        int ilo = d - ly + 1;
        if (ilo < 0) ilo = 0;
        const int ilo2 = floordiv2(d - max_len) + 1;
        if (ilo2 > ilo) ilo = ilo2;
        int ihi = d < lx - 1 ? d : lx - 1;
        const int ihi2 = floordiv2(d + min_len + w - 1);
        if (ihi2 < ihi) ihi = ihi2;
        const int s = ilo - 1 > 0 ? ilo - 1 : 0;
        const int e = ihi + 1 < lx - 1 ? ihi + 1 : lx - 1;
        for (int i = s + tid; i <= e; i += blockDim.x) {
            T c = INF;
            if (i >= ilo && i <= ihi) {
                const int j = d - i;
                const T v = X[i] - Y[j];
                T wt = (T)1.0;
                if (weights) wt = (i == 0 && j > 0) ? weights[j - 1] : weights[i > j ? i - j : j - i];
                if (i == 0 && j == 0) {
                    c = v * v * wt;
                } else if constexpr (AD) {
                    const T up = i > 0 ? d1[i - 1] : INF;
                    const T left = j > 0 ? d1[i] : INF;
                    const T dg = (i > 0 && j > 0) ? d2[i - 1] : INF;
                    c = i > 0 ? cmin(cmin(up + pen, left + pen), dg) + v * v : left + v * v;
                } else {
                    const T up = i > 0 ? d1[i - 1] : INF;
                    const T left = j > 0 ? d1[i] : INF;
                    const T dg = (i > 0 && j > 0) ? d2[i - 1] : INF;
                    c = cmin(cmin(up, left), dg) + v * v * wt;
                }
            }
            d0[i] = c;
        }
        __syncthreads();
        T *t = d2;
        d2 = d1;
        d1 = d0;
        d0 = t;
    }
    const T res = d1[lx - 1];
    __syncthreads();
    return res;
}

}  // namespace wbx
