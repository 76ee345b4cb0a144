// path.hip -- float64 optimal warping paths and the majorize-minimize
// barycenter update.
//
// k_path_moves: one pair per 64-lane block.  The anti-diagonal sweep of
//   block_dtw computes the cells of _dtw_alignment (distance/_elastic.pyx:
//   1011-1073; for DTW the same cell arithmetic as dtw_distance) with band
//   width w (and, for wdtw_alignment, weights w[|i - j|]), and stores for
//   every band cell the move dtw_mapping
//   (distance/dtw.py:362-427) takes out of it: argmin over (diag, up, left)
//   of the neighbours' costs, first wins, +inf outside the band or matrix.
// k_path_walk: one pair per thread walks the moves from (lx-1, ly-1) to
//   (0, 0); a monotone path covers a contiguous column range of each row, so
//   the path is returned as rows[p][2 i] = first column, rows[p][2 i + 1] =
//   last column.
// k_mm_update: _mm_dtw_average's accumulation (distance/dtw.py:650-666):
//   thread m adds, for series i = 0 .. n-1 in order and the path columns x of
//   row m in increasing order, V[m] += w and z[m] += X[i, x] * w, then
//   mean[m] = z[m] / V[m] -- the reference's order of additions exactly.
// k_ssg_update: one step of _ssg_dtw_average (distance/dtw.py:597-647) for
//   one series along its path.
#pragma clang fp contract(off)

#include "block_dtw.h"

namespace wbx {

namespace {

// Move codes: 0 diag, 1 up, 2 left.
__device__ __forceinline__ unsigned char move_of(double diag, double up, double left) {
    if (diag <= up && diag <= left) return 0;
    return up <= left ? 1 : 2;
}

__global__ __launch_bounds__(64) void k_path_moves(const double *__restrict__ X, int64_t xstride,
                                                   const double *__restrict__ Y, int64_t ystride,
                                                   int64_t n, int lx, int ly, int w, int B,
                                                   int off, int cap,
                                                   const double *__restrict__ weights,
                                                   unsigned char *__restrict__ moves) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *d0 = reinterpret_cast<double *>(smem_raw);
    double *d1 = d0 + cap, *d2 = d0 + 2 * cap;
    const int64_t p = blockIdx.x;
    if (p >= n) return;
    const double *__restrict__ x = X + p * xstride;
    const double *__restrict__ y = Y + p * ystride;
    unsigned char *__restrict__ mv = moves + p * (int64_t)lx * B;
    const int min_len = lx > ly ? lx - ly : 0;
    const int max_len = (ly > lx ? ly - lx : 0) + w;
    const double INF = WBX_INF_D;
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
            double c = INF;
            if (i >= ilo && i <= ihi) {
                const int j = d - i;
                const double v = x[i] - y[j];
                // _dtw_alignment weighs every cell, row 0 included, by
                // w[|i - j|] (1 without weights: v * v * 1 == v * v)
                const double wt = weights ? weights[i > j ? i - j : j - i] : 1.0;
                if (i == 0 && j == 0) {
                    c = v * v * wt;
                } else {
                    const double up = i > 0 ? d1[i - 1] : INF;
                    const double left = j > 0 ? d1[i] : INF;
                    const double dg = (i > 0 && j > 0) ? d2[i - 1] : INF;
                    c = cmin(cmin(up, left), dg) + v * v * wt;
                    mv[(int64_t)i * B + (j - i + off)] = move_of(dg, up, left);
                }
            }
            d0[i] = c;
        }
        __syncthreads();
        double *t = d2;
        d2 = d1;
        d1 = d0;
        d0 = t;
    }
}

// One pair per thread: the walk from (lx-1, ly-1) to (0, 0).  On the first
// row only left moves exist and on the first column only up moves (all three
// options +inf happen only when costs overflow; the walk then stays inside
// the matrix).
__global__ __launch_bounds__(256) void k_path_walk(int64_t n, int lx, int ly, int B, int off,
                                                   const unsigned char *__restrict__ moves,
                                                   int32_t *__restrict__ rows) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const unsigned char *__restrict__ mv = moves + p * (int64_t)lx * B;
    int32_t *__restrict__ rw = rows + p * (int64_t)lx * 2;
    int i = lx - 1, j = ly - 1;
    rw[2 * i + 1] = j;
    while (i > 0 || j > 0) {
        unsigned char m = mv[(int64_t)i * B + (j - i + off)];
        if (i == 0) m = 2;
        else if (j == 0) m = 1;
        if (m == 2) {
            --j;
            continue;
        }
        rw[2 * i] = j;
        --i;
        if (m == 0) --j;
        rw[2 * i + 1] = j;
    }
    rw[0] = 0;
}

__global__ __launch_bounds__(256) void k_mm_update(const int32_t *__restrict__ rows,
                                                   const double *__restrict__ X, int64_t xstride,
                                                   int64_t n, int lm,
                                                   const double *__restrict__ sample_weight,
                                                   double *__restrict__ mean) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m >= lm) return;
    double z = 0.0, V = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const int32_t lo = rows[(i * lm + m) * 2], hi = rows[(i * lm + m) * 2 + 1];
        const double *__restrict__ xi = X + i * xstride;
        const double wgt = sample_weight ? sample_weight[i] : 1.0;
        for (int32_t c = lo; c <= hi; ++c) {
            V += wgt;
            z += xi[c] * wgt;
        }
    }
    mean[m] = z / V;
}

// _ssg_dtw_average's step (distance/dtw.py:618-633): thread m sums, over the
// path columns x of row m in increasing order, z += mean[m] - x[x] * w, then
// mean[m] -= lr * z (every z from the mean before the step).
__global__ __launch_bounds__(256) void k_ssg_update(const int32_t *__restrict__ rows,
                                                    const double *__restrict__ x, double w,
                                                    double lr, int lm, double *__restrict__ mean) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m >= lm) return;
    const int32_t lo = rows[2 * m], hi = rows[2 * m + 1];
    const double mm = mean[m];
    double z = 0.0;
    for (int32_t c = lo; c <= hi; ++c) z += mm - x[c] * w;
    mean[m] = mm - lr * z;
}

}  // namespace

int launch_ssg_update(const int32_t *rows, const double *x, double w, double lr, int lm,
                      double *mean, hipStream_t st) {
    if (lm <= 0) return WBX_OK;
    hipLaunchKernelGGL(k_ssg_update, dim3((unsigned)((lm + 255) / 256)), dim3(256), 0, st, rows, x,
                       w, lr, lm, mean);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

int launch_path_rows(const double *X, int64_t xstride, const double *Y, int64_t ystride,
                     int64_t n, int lx, int ly, int w, const double *weights,
                     unsigned char *moves, int32_t *rows, hipStream_t st) {
    if (n <= 0) return WBX_OK;
    const Band bd = make_band(lx, ly, 0.0, w);
    const int cap = lx;
    const size_t lds = (size_t)3 * (size_t)cap * sizeof(double);
    if (lds > 64 * 1024) {
        if (lds > 160 * 1024) return fail(WBX_EUNSUPPORTED, "path: series of %d too long", lx);
        WBX_HIP_CHECK(hipFuncSetAttribute((const void *)k_path_moves,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    if (n > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "path: too many pairs");
    hipLaunchKernelGGL(k_path_moves, dim3((unsigned)n), dim3(64), lds, st, X, xstride, Y, ystride,
                       n, lx, ly, w, bd.B, bd.off, cap, weights, moves);
    WBX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_path_walk, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, lx, ly,
                       bd.B, bd.off, moves, rows);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

int launch_mm_update(const int32_t *rows, const double *X, int64_t xstride, int64_t n, int lm,
                     const double *sample_weight, double *mean, hipStream_t st) {
    if (lm <= 0) return WBX_OK;
    hipLaunchKernelGGL(k_mm_update, dim3((unsigned)((lm + 255) / 256)), dim3(256), 0, st, rows, X,
                       xstride, n, lm, sample_weight, mean);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

}  // namespace wbx
