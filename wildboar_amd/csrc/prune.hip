// prune.hip -- argmin_distance with a lower bound: which pairs need a DTW.
//
// The reference (distance/_cdistance.pyx:1329-1340) walks y in index order
// with a threshold T (+inf until the heap holds k entries, then the heap's
// largest value) and skips pair j when lower_bound[i, j] >= T; otherwise it
// computes the distance (abandoning at T) and pushes it when it is below T.
// A skipped or abandoned pair is never among the k smallest values seen so
// far, so T before pair j is the k-th smallest distance over all pairs
// j' < j.  The reference therefore evaluates exactly E = {j : lb[j] < T(j)}.
//
// Let S be a set of pairs with exactly known distances v on which the bound
// holds (lb[s] <= v[s]), and U_S(j) the k-th smallest known value over
// s < j (+inf below k of them).  Then U_S(j) >= T(j): the known values the
// reference evaluated are among its own (at most k - 1 of them below T(j)),
// and a pair s it skipped has v[s] >= lb[s] >= T(s) >= T(j) (T only falls).
// So C_S = {j : lb[j] < U_S(j)} contains E, whatever the bound is elsewhere.
// The library computes S = the first k pairs and the kSeed pairs of smallest
// bound of every row, checks the bound on S (a row where it fails lists every
// pair with a finite bound), then the distances of C_S outside S: every pair
// the reference evaluates has its exact value, and the heap replay
// (k_argmin_replay_wave) applies the reference's own skip rule to the rest
// (their entries stay NaN and are never read as candidates).
#include "wbx_internal.h"

namespace wbx {

namespace {

constexpr int kSeed = 16;  // smallest-bound seeds per row
constexpr int kKMax = 16;  // largest k served

__device__ __forceinline__ void push_pair(int32_t *pa, int32_t *pb, unsigned long long *cnt,
                                          int64_t i, int64_t j) {
    const unsigned long long q = atomicAdd(cnt, 1ull);
    pa[q] = (int32_t)i;
    pb[q] = (int32_t)j;
}

// One thread per row: the first k pairs and the kSeed pairs of smallest bound
// (bound < +inf; ties by index).
__global__ __launch_bounds__(64) void k_prune_seed(const double *__restrict__ lb, int64_t nx,
                                                   int64_t ny, int64_t k, int32_t *__restrict__ pa,
                                                   int32_t *__restrict__ pb,
                                                   unsigned long long *__restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (i >= nx) return;
    const double *row = lb + i * ny;
    double sv[kSeed];
    int64_t sj[kSeed];
#pragma unroll
    for (int q = 0; q < kSeed; ++q) {
        sv[q] = WBX_INF_D;
        sj[q] = -1;
    }
    for (int64_t j = k; j < ny; ++j) {
        double v = row[j];
        int64_t jj = j;
        if (!(v < sv[kSeed - 1])) continue;
#pragma unroll
        for (int q = 0; q < kSeed; ++q) {  // sorted insertion
            if (v < sv[q]) {
                const double tv = sv[q];
                const int64_t tj = sj[q];
                sv[q] = v;
                sj[q] = jj;
                v = tv;
                jj = tj;
            }
        }
    }
    for (int64_t j = 0; j < k; ++j)
        if (row[j] < WBX_INF_D) push_pair(pa, pb, cnt, i, j);
#pragma unroll
    for (int q = 0; q < kSeed; ++q)
        if (sj[q] >= 0) push_pair(pa, pb, cnt, i, sj[q]);
}

// One thread per row: with D holding the known distances (NaN elsewhere),
// every pair with lb[j] < U(j) whose distance is not known is listed.
__global__ __launch_bounds__(64) void k_prune_scan(const double *__restrict__ D,
                                                   const double *__restrict__ lb, int64_t nx,
                                                   int64_t ny, int k, int32_t *__restrict__ pa,
                                                   int32_t *__restrict__ pb,
                                                   unsigned long long *__restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (i >= nx) return;
    const double *drow = D + i * ny;
    const double *lrow = lb + i * ny;
    // the bound must hold on the known pairs; else every pair with a finite
    // bound is listed (the replay is then exact for any bound)
    bool holds = true;
    for (int64_t j = 0; j < ny; ++j) {
        const double v = drow[j];
        if (v == v && lrow[j] > v) holds = false;
    }
    double top[kKMax];  // the k smallest known values so far, ascending
#pragma unroll
    for (int q = 0; q < kKMax; ++q) top[q] = WBX_INF_D;
    double U = WBX_INF_D;
    for (int64_t j = 0; j < ny; ++j) {
        const double v = drow[j];
        const bool known = v == v;
        if (!known && lrow[j] < (holds ? U : WBX_INF_D)) push_pair(pa, pb, cnt, i, j);
        if (known && v < U) {  // U = top[k - 1], or +inf below k values
            double c = v;
#pragma unroll
            for (int q = 0; q < kKMax; ++q) {
                if (c < top[q]) {
                    const double t = top[q];
                    top[q] = c;
                    c = t;
                }
            }
#pragma unroll
            for (int q = 0; q < kKMax; ++q)
                if (q == k - 1) U = top[q];
        }
    }
}

}  // namespace

int prune_k_max() { return kKMax; }

int launch_prune_seed(const double *lb, int64_t nx, int64_t ny, int64_t k, int32_t *pa,
                      int32_t *pb, unsigned long long *cnt, hipStream_t st) {
    if (k > kKMax) return fail(WBX_EINVAL, "bound pruning serves k <= %d", kKMax);
    hipLaunchKernelGGL(k_prune_seed, dim3((unsigned)((nx + 63) / 64)), dim3(64), 0, st, lb, nx, ny,
                       k, pa, pb, cnt);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

int launch_prune_scan(const double *D, const double *lb, int64_t nx, int64_t ny, int64_t k,
                      int32_t *pa, int32_t *pb, unsigned long long *cnt, hipStream_t st) {
    if (k > kKMax) return fail(WBX_EINVAL, "bound pruning serves k <= %d", kKMax);
    hipLaunchKernelGGL(k_prune_scan, dim3((unsigned)((nx + 63) / 64)), dim3(64), 0, st, D, lb, nx,
                       ny, (int)k, pa, pb, cnt);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

}  // namespace wbx
