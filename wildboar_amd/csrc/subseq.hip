// subseq.hip -- subsequence scans (minimum over windows + first argmin).
//
// _PairwiseSubsequenceDistance (distance/_cdistance.pyx:1001-1015) writes
// dist[i, j] / idx[i, j] for sample i and shapelet j.  Per (i, j):
//   dtw / wdtw : dtw_subsequence_distance (distance/_elastic.pyx:622-655):
//                for window s: cost = dtw_distance(S, L, x + s, L, w), w =
//                _compute_r(L, r); keep the first strictly smaller squared
//                cost; return sqrt(best).
//   euclidean  : euclidean_distance (distance/_metric.pyx:654-683) with the
//                same first-strict-minimum rule on squared sums.
// The reference's early abandoning never changes a result (a window is
// abandoned only when its cost already exceeds the best), so the kernels here
// evaluate windows fully where that is simpler.
//
// k_subseq_exact_dtw: one (i, j) per 64-lane block, windows in order, each
// window by the plain anti-diagonal DP (bit-identical in FP64).
#pragma clang fp contract(off)

#include "block_dtw.h"

namespace wbx {

// band_add / AD: as k_subseq_lane64 (derivative patterns, amercing).
template <typename T, bool AD>
__global__ __launch_bounds__(64) void k_subseq_exact_dtw(Shapelets sh, Series<T> xs, int64_t n,
                                                         double r, const T *__restrict__ weights,
                                                         int cap, int lane_cap, int band_add,
                                                         T pen, double *__restrict__ dist,
                                                         int64_t *__restrict__ idx) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *buf = reinterpret_cast<T *>(smem_raw);
    const int64_t q = (int64_t)blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
    if (q >= n * sh.n) return;
    const int64_t i = q / sh.n, j = q - i * sh.n;
    const T *S = reinterpret_cast<const T *>(sh.S) + sh.off[j];
    const int L = (int)sh.len[j];
    const int t = xs.len;
    const T *x = xs.row(i);
    if (L < 1) return;  // handled by the caller
    const int w = band_w(L + band_add, L + band_add, r);
    if ((2 * w - 1 + 7) / 8 * 8 <= lane_cap) return;  // a lane64 launch takes this task
    T best = (T)WBX_INF_D;
    int64_t best_i = 0;
    for (int s = 0; s <= t - L; ++s) {
        const T c = block_dtw<T, AD>(S, L, x + s, L, w, weights, buf, buf + cap, buf + 2 * cap,
                                     pen);
        if (c < best) {
            best = c;
            best_i = s;
        }
    }
    if (threadIdx.x == 0) {
        dist[q] = sqrt((double)best);
        idx[q] = best_i;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_subseq_exact_euclid(Shapelets sh, Series<T> xs,
                                                             int64_t n, double *__restrict__ dist,
                                                             int64_t *__restrict__ idx) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n * sh.n) return;
    const int64_t i = q / sh.n, j = q - i * sh.n;
    const T *S = reinterpret_cast<const T *>(sh.S) + sh.off[j];
    const int L = (int)sh.len[j];
    const int t = xs.len;
    const T *x = xs.row(i);
    T min_dist = (T)WBX_INF_D;
    int64_t best_i = 0;
    for (int s = 0; s < t - L + 1; ++s) {
        T d = (T)0;
        for (int k = 0; k < L; ++k) {
            if (d >= min_dist) break;
            T v = x[s + k];
            v -= S[k];
            d += v * v;
        }
        if (d < min_dist) {
            min_dist = d;
            best_i = s;
        }
    }
    dist[q] = sqrt((double)min_dist);
    idx[q] = best_i;
}

// dist = 0, idx = 0 for the patterns of length < 1 (derivative patterns of
// series shorter than 3: DerivativeDtwSubsequenceMetric returns 0 there,
// distance/_elastic.pyx:792-793, and leaves the index unset).
__global__ __launch_bounds__(256) void k_subseq_short(const int64_t *__restrict__ len,
                                                      int64_t n_s, int64_t n,
                                                      double *__restrict__ dist,
                                                      int64_t *__restrict__ idx) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n * n_s) return;
    if (len[q % n_s] < 1) {
        dist[q] = 0.0;
        idx[q] = 0;
    }
}

int launch_subseq_short(const int64_t *len, int64_t n_s, int64_t n, double *dist, int64_t *idx,
                        hipStream_t st) {
    const int64_t total = n * n_s;
    if (total <= 0) return WBX_OK;
    if ((total + 255) / 256 > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "too many tasks");
    hipLaunchKernelGGL(k_subseq_short, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       len, n_s, n, dist, idx);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

template <typename T>
int launch_subseq_exact(int kind, double r, Shapelets sh, int max_len, Series<T> xs, int64_t n,
                        const T *weights, double *dist, int64_t *idx, hipStream_t st,
                        int lane_cap, int band_add, const double *penalty) {
    const int64_t nq = n * sh.n;
    if (nq <= 0) return WBX_OK;
    if (kind == WBX_EUCLIDEAN) {
        hipLaunchKernelGGL(k_subseq_exact_euclid<T>, dim3((unsigned)((nq + 255) / 256)), dim3(256),
                           0, st, sh, xs, n, dist, idx);
    } else {
        const size_t lds = (size_t)3 * (size_t)max_len * sizeof(T);
        if (lds > 64 * 1024) {
            if (lds > 160 * 1024)
                return fail(WBX_EUNSUPPORTED, "shapelet of length %d exceeds LDS", max_len);
            WBX_HIP_CHECK(hipFuncSetAttribute((const void *)k_subseq_exact_dtw<T, false>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)lds));
            WBX_HIP_CHECK(hipFuncSetAttribute((const void *)k_subseq_exact_dtw<T, true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)lds));
        }
        dim3 grid((unsigned)(nq < 65536 ? nq : 65536), (unsigned)((nq + 65535) / 65536));
        if (penalty)
            hipLaunchKernelGGL((k_subseq_exact_dtw<T, true>), grid, dim3(64), lds, st, sh, xs, n,
                               r, weights, max_len, lane_cap, band_add, (T)*penalty, dist, idx);
        else
            hipLaunchKernelGGL((k_subseq_exact_dtw<T, false>), grid, dim3(64), lds, st, sh, xs, n,
                               r, weights, max_len, lane_cap, band_add, (T)0, dist, idx);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

template int launch_subseq_exact<double>(int, double, Shapelets, int, Series<double>, int64_t,
                                         const double *, double *, int64_t *, hipStream_t, int,
                                         int, const double *);
template int launch_subseq_exact<float>(int, double, Shapelets, int, Series<float>, int64_t,
                                        const float *, double *, int64_t *, hipStream_t, int, int,
                                        const double *);

}  // namespace wbx
