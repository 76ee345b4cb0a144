// subseq_match.hip -- distance profiles: the squared cost of every window of a
// series against a pattern, exact in float64 (the reference's cell order,
// contraction off).  They serve
//   argmin_subsequence_distance: _ArgminSubsequenceDistance
//       (distance/_cdistance.pyx:1408-1434), a heap over the windows in
//       order, replayed by k_argmin_replay_wave on the square roots;
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
// Euclidean: lane g owns the four windows s = 4 g .. 4 g + 3 and keeps
// x[4 g + u .. 4 g + u + 8) in registers, so each pattern value (uniform
// across the wave) meets four independent sums; every sum runs in the
// reference's order (v = x[s + u] - S[u], cost += v * v).  Four steps use
// static register indices, then the window moves by four with two 16-byte
// loads.  Reads past the row end are clamped (they only feed windows past
// the last one).
typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));

__device__ __forceinline__ void ld2(const double *__restrict__ x, int e, int t, double &a,
                                    double &b) {
    if (e + 1 < t) {
        const d2u v = *(const d2u *)(x + e);
        a = v[0];
        b = v[1];
    } else {
        a = x[e < t ? e : t - 1];
        b = x[t - 1];
    }
}

__global__ __launch_bounds__(256) void k_profile_euclid(Shapelets sh, int paired,
                                                        Series<double> xs, int64_t n, int64_t nwp,
                                                        int nblk, double *__restrict__ prof) {
    const int64_t q = blockIdx.x / nblk;
    const int blk = (int)(blockIdx.x - q * nblk);
    const int w0 = (blk * 256 + (int)threadIdx.x) * 4;
    if (w0 >= nwp) return;
    const int64_t j = paired ? q : 0;
    const double *__restrict__ S = (const double *)sh.S + sh.off[j];
    const int L = (int)sh.len[j];
    const int t = xs.len;
    const double *__restrict__ x = xs.row(q);
    double xw[8];
#pragma unroll
    for (int m = 0; m < 8; m += 2) ld2(x, w0 + m, t, xw[m], xw[m + 1]);
    double c[4] = {0.0, 0.0, 0.0, 0.0};
    int u = 0;
    for (; u + 4 <= L; u += 4) {
#pragma unroll
        for (int uu = 0; uu < 4; ++uu) {
            const double sv = S[u + uu];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                double v = xw[k + uu];
                v -= sv;
                c[k] += v * v;
            }
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) xw[m] = xw[m + 4];
        ld2(x, w0 + u + 8, t, xw[4], xw[5]);
        ld2(x, w0 + u + 10, t, xw[6], xw[7]);
    }
    for (; u < L; ++u) {
        const double sv = S[u];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double v = xw[k];
            v -= sv;
            c[k] += v * v;
        }
#pragma unroll
        for (int m = 0; m < 7; ++m) xw[m] = xw[m + 1];
        const int e = w0 + u + 8;
        xw[7] = x[e < t ? e : t - 1];
    }
    const int nwin = t - L + 1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int s = w0 + k;
        if (s < nwp) prof[q * nwp + s] = s < nwin ? c[k] : WBX_INF_D;
    }
}

// DTW / WDTW: one 64-lane block per (task, chunk of windows), windows in
// order, each by the anti-diagonal DP of block_dtw (bit-identical in FP64).
__global__ __launch_bounds__(64) void k_profile_dtw(Shapelets sh, int paired, Series<double> xs,
                                                    int64_t n, int64_t nwp, int chunks, int chunk,
                                                    double r, const double *__restrict__ weights,
                                                    int cap, int lane_cap,
                                                    double *__restrict__ prof) {
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
    if ((2 * w - 1 + 7) / 8 * 8 <= lane_cap) return;  // a lane64 launch writes this row
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

int launch_profile(int kind, double r, Shapelets sh, int paired, int min_len, int max_len,
                   Series<double> xs, int64_t n, int64_t nwp, const double *weights, bool lane_ok,
                   double *prof, hipStream_t st) {
    const int64_t total = n * nwp;
    if (total <= 0) return WBX_OK;
    if (kind == WBX_EUCLIDEAN) {
        const int nblk = (int)((nwp + 1023) / 1024);
        if (n * nblk > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "profile grid too large");
        hipLaunchKernelGGL(k_profile_euclid, dim3((unsigned)(n * nblk)), dim3(256), 0, st, sh,
                           paired, xs, n, nwp, nblk, prof);
    } else {
        const bool wt = weights != nullptr;
        int lane_cap = 0;
        if (lane_ok) {
            const int rc = launch_profile_lane64(r, sh, paired, min_len, max_len, xs, n, nwp,
                                                 weights, prof, st);
            if (rc) return rc;
            lane_cap = lane64_cap(wt);
            // every row fits a lane class: nothing left for the block kernel
            if (lane64_class(2 * band_w(max_len, max_len, r) - 1, wt)) return WBX_OK;
        }
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
                           kChunk, r, weights, max_len, lane_cap, prof);
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
