// rerank.hip -- certified selection of the pairs an FP32 result cannot decide.
//
// An FP32 cell is c = fl(fma(d, d, m)) with d = fl(x - y) (or fl(fl(d*d)*w)+m
// in the plain FP32 kernel).  All terms are non-negative and min is exact, so
// by induction over the DP every computed cost satisfies
//     (1-u)^K C - A  <=  c  <=  (1+u)^K C + A,    K = lx + ly + 2 (+2 WDTW)
// with u = 2^-24, C the exact real cost and A = P * 4 u M^2 (1+u)^K the
// contribution of rounding float64 inputs to float32 (M = max |input|,
// P = lx + ly - 1 path cells; A = 0 when the inputs already are float32).
// The reference's float64 value obeys the same bound with u = 2^-53.  This
// gives for every pair an interval [lo, hi] that contains the reference's
// sqrt'ed float64 distance.
//
// argmin (heap replay of _ArgMinBatch, distance/_cdistance.pyx:1318-1345):
// pair j can be pushed only if d_j < T_j, where T_j is the k-th smallest of
// d_0..d_{j-1}; T_j <= the k-th smallest hi over j' < j.  So every j with
// lo_j < that running k-th smallest hi is a candidate and nothing else can be
// pushed.  Candidates are re-evaluated exactly in FP64 and the replay runs on
// exact values; non-candidates are +inf and are never pushed.
#include "wbx_internal.h"

namespace wbx {

ErrBound make_err_bound(int lx, int ly, int weighted, double absmax, int inputs_exact) {
    const double u = 5.9604644775390625e-08;  // 2^-24
    const double u64 = 1.1102230246251565e-16;
    const int P = lx + ly - 1;
    const double K = (double)P + 3.0 + (weighted ? 2.0 : 0.0);
    ErrBound eb;
    eb.e_up = exp(K * log1p(u)) * (1.0 + 1e-12);
    eb.e_dn = exp(K * log1p(-u)) * (1.0 - 1e-12);
    eb.A = inputs_exact ? 0.0 : (double)P * 4.0 * u * absmax * absmax * eb.e_up * (1.0 + 1e-12);
    eb.e64 = (K + 8.0) * u64 * 1.01;
    return eb;
}

__device__ __forceinline__ void dist_bounds(float c, const ErrBound &eb, double &lo, double &hi) {
    if (!(c < WBX_INF_F)) {
        lo = 0.0;
        hi = WBX_INF_D;
        return;
    }
    const double cd = (double)c;
    double clo = cd - eb.A;
    clo = clo > 0.0 ? clo / eb.e_up : 0.0;
    const double chi = (cd + eb.A) / eb.e_dn;
    lo = sqrt(clo * (1.0 - eb.e64)) * (1.0 - 4.4e-16);
    hi = sqrt(chi * (1.0 + eb.e64)) * (1.0 + 4.4e-16);
}

constexpr int KMAX = 32;

// One lane per x row.  write == 0: count candidates; write == 1: emit them at
// offs[i] into (pa, pb).
__global__ __launch_bounds__(64) void k_argmin_candidates(const float *__restrict__ sq,
                                                          int64_t nx, int64_t ny, int k,
                                                          ErrBound eb, int write,
                                                          int64_t *__restrict__ counts,
                                                          const int64_t *__restrict__ offs,
                                                          int32_t *__restrict__ pa,
                                                          int32_t *__restrict__ pb) {
    const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (i >= nx) return;
    double top[KMAX];
    int nf = 0;
    int64_t cnt = 0;
    const int64_t base = write ? offs[i] : 0;
    for (int64_t j = 0; j < ny; ++j) {
        double lo, hi;
        dist_bounds(sq[i * ny + j], eb, lo, hi);
        const double tup = nf < k ? WBX_INF_D : top[k - 1];
        if (lo < tup) {
            if (write) {
                pa[base + cnt] = (int32_t)i;
                pb[base + cnt] = (int32_t)j;
            }
            ++cnt;
        }
        // keep the k smallest hi values, ascending
        if (nf < k || hi < top[k - 1]) {
            int q = nf < k ? nf++ : k - 1;
            while (q > 0 && top[q - 1] > hi) {
                top[q] = top[q - 1];
                --q;
            }
            top[q] = hi;
        }
    }
    if (!write) counts[i] = cnt;
}

// Exclusive scan of counts (nx values) into offs[0..nx], one block.
__global__ __launch_bounds__(1024) void k_scan_counts(const int64_t *__restrict__ counts,
                                                      int64_t nx, int64_t *__restrict__ offs) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (nx + 1023) / 1024;
    const int64_t s = (int64_t)t * per, e = s + per < nx ? s + per : nx;
    int64_t acc = 0;
    for (int64_t q = s; q < e; ++q) acc += counts[q];
    part[t] = acc;
    __syncthreads();
    if (t == 0) {
        int64_t run = 0;
        for (int q = 0; q < 1024; ++q) {
            int64_t v = part[q];
            part[q] = run;
            run += v;
        }
        offs[nx] = run;
    }
    __syncthreads();
    acc = part[t];
    for (int64_t q = s; q < e; ++q) {
        offs[q] = acc;
        acc += counts[q];
    }
}

int launch_argmin_candidates(const float *sq, int64_t nx, int64_t ny, int k, const ErrBound &eb,
                             int64_t *counts, int64_t *offs, int32_t *pa, int32_t *pb,
                             int64_t *total_host, hipStream_t st) {
    if (k > KMAX) return fail(WBX_EUNSUPPORTED, "certified argmin supports k <= %d", KMAX);
    const dim3 grid((unsigned)((nx + 63) / 64));
    hipLaunchKernelGGL(k_argmin_candidates, grid, dim3(64), 0, st, sq, nx, ny, k, eb, 0, counts,
                       offs, pa, pb);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, counts, nx, offs);
    hipLaunchKernelGGL(k_argmin_candidates, grid, dim3(64), 0, st, sq, nx, ny, k, eb, 1, counts,
                       offs, pa, pb);
    WBX_HIP_CHECK(hipGetLastError());
    WBX_HIP_CHECK(hipMemcpyAsync(total_host, offs + nx, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    return WBX_OK;
}

}  // namespace wbx
