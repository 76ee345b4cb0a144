// subseq_match.hip -- distance profiles: the squared cost of every window of a
// series against a pattern, exact in float64 (the reference's cell order,
// contraction off).  They serve
//   argmin_subsequence_distance: _ArgminSubsequenceDistance
//       (distance/_cdistance.pyx:1408-1434), a heap over the windows in
//       order, replayed by k_argmin_replay on the square roots;
//   subsequence_match: dtw_subsequence_matches (distance/_elastic.pyx:658-698)
//       and euclidean_distance_matches (distance/_metric.pyx:762-795), the
//       windows whose cost is <= threshold^2.
// Early abandoning in those loops never changes which windows qualify (a
// window is abandoned only once its partial cost exceeds the bound), so every
// window is evaluated in full.
#pragma clang fp contract(off)

#include "block_dtw.h"

namespace wbx {

// Task q: pattern (paired ? q : 0) against series q.  prof[q * nwp + s] is
// the squared cost of window s, +inf for windows past the pattern's last.
// Euclidean: one thread per window, the sum in the reference's order
// (v = x[s + u] - S[u], cost += v * v).
__global__ __launch_bounds__(256) void k_profile_euclid(Shapelets sh, int paired,
                                                        Series<double> xs, int64_t n, int64_t nwp,
                                                        double *__restrict__ prof) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= n * nwp) return;
    const int64_t q = g / nwp, s = g - q * nwp;
    const int64_t j = paired ? q : 0;
    const double *S = (const double *)sh.S + sh.off[j];
    const int64_t L = sh.len[j];
    double c = WBX_INF_D;
    if (s + L <= xs.len) {
        const double *x = xs.row(q) + s;
        c = 0.0;
        for (int64_t u = 0; u < L; ++u) {
            double v = x[u];
            v -= S[u];
            c += v * v;
        }
    }
    prof[g] = c;
}

// DTW / WDTW: one 64-lane block per (task, chunk of windows), windows in
// order, each by the anti-diagonal DP of block_dtw (bit-identical in FP64).
__global__ __launch_bounds__(64) void k_profile_dtw(Shapelets sh, int paired, Series<double> xs,
                                                    int64_t n, int64_t nwp, int chunks, int chunk,
                                                    double r, const double *__restrict__ weights,
                                                    int cap, double *__restrict__ prof) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *buf = reinterpret_cast<double *>(smem_raw);
    const int64_t b = (int64_t)blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
    if (b >= n * chunks) return;
    const int64_t q = b / chunks;
    const int c = (int)(b - q * chunks);
    const int64_t j = paired ? q : 0;
    const double *S = (const double *)sh.S + sh.off[j];
    const int L = (int)sh.len[j];
    const double *x = xs.row(q);
    const int w = band_w(L, L, r);
    const int64_t s0 = (int64_t)c * chunk;
    const int64_t s1 = s0 + chunk < nwp ? s0 + chunk : nwp;
    for (int64_t s = s0; s < s1; ++s) {
        double v = WBX_INF_D;
        if (s + L <= xs.len)
            v = block_dtw<double>(S, L, x + s, L, w, weights, buf, buf + cap, buf + 2 * cap);
        if (threadIdx.x == 0) prof[q * nwp + s] = v;
    }
}

__global__ __launch_bounds__(256) void k_sqrt_inplace(double *__restrict__ p, int64_t n) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g < n) p[g] = sqrt(p[g]);
}

int launch_profile(int kind, double r, Shapelets sh, int paired, int max_len, Series<double> xs,
                   int64_t n, int64_t nwp, const double *weights, double *prof, hipStream_t st) {
    const int64_t total = n * nwp;
    if (total <= 0) return WBX_OK;
    if (kind == WBX_EUCLIDEAN) {
        hipLaunchKernelGGL(k_profile_euclid, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                           st, sh, paired, xs, n, nwp, prof);
    } else {
        const size_t lds = (size_t)3 * (size_t)max_len * sizeof(double);
        if (lds > 64 * 1024) {
            if (lds > 160 * 1024)
                return fail(WBX_EUNSUPPORTED, "pattern of length %d exceeds LDS", max_len);
            WBX_HIP_CHECK(hipFuncSetAttribute((const void *)k_profile_dtw,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)lds));
        }
        constexpr int kChunk = 16;
        const int chunks = (int)((nwp + kChunk - 1) / kChunk);
        const int64_t nb = n * chunks;
        dim3 grid((unsigned)(nb < 65536 ? nb : 65536), (unsigned)((nb + 65535) / 65536));
        hipLaunchKernelGGL(k_profile_dtw, grid, dim3(64), lds, st, sh, paired, xs, n, nwp, chunks,
                           kChunk, r, weights, max_len, prof);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

int launch_sqrt_inplace(double *p, int64_t n, hipStream_t st) {
    if (n <= 0) return WBX_OK;
    hipLaunchKernelGGL(k_sqrt_inplace, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

}  // namespace wbx
