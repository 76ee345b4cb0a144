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

// Lexicographic minimum of (v, j) over the 64 lanes (every lane gets it).
__device__ __forceinline__ void wave_min_vj(double &v, int &j) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const double ov = __shfl_xor(v, m);
        const int oj = __shfl_xor(j, m);
        if (ov < v || (ov == v && oj < j)) {
            v = ov;
            j = oj;
        }
    }
}

// One wave per row: the first k pairs (bound < +inf) and the kSeed pairs of
// smallest bound among j >= k (ties by index).  Each lane keeps the kSeed
// smallest of its columns k + lane, k + lane + 64, ... (sorted insertion in
// increasing j, so an equal value never displaces an earlier one); kSeed
// rounds of a wave-wide (value, index) minimum then take the row's kSeed
// smallest -- the set a scan of the row in index order keeps.
__global__ __launch_bounds__(64) void k_prune_seed(const double *__restrict__ lb, int64_t nx,
                                                   int64_t ny, int64_t k, int32_t *__restrict__ pa,
                                                   int32_t *__restrict__ pb,
                                                   unsigned long long *__restrict__ cnt) {
    const int64_t i = blockIdx.x;
    if (i >= nx) return;
    const int lane = threadIdx.x;
    const double *row = lb + i * ny;
    for (int64_t j = lane; j < k && j < ny; j += 64)
        if (row[j] < WBX_INF_D) push_pair(pa, pb, cnt, i, j);
    constexpr int kNone = 0x7fffffff;
    double sv[kSeed];
    int sj[kSeed];
#pragma unroll
    for (int q = 0; q < kSeed; ++q) {
        sv[q] = WBX_INF_D;
        sj[q] = kNone;
    }
    for (int64_t j = k + lane; j < ny; j += 64) {
        double v = row[j];
        int jj = (int)j;
        if (!(v < sv[kSeed - 1])) continue;
#pragma unroll
        for (int q = 0; q < kSeed; ++q) {  // sorted insertion
            if (v < sv[q]) {
                const double tv = sv[q];
                const int tj = sj[q];
                sv[q] = v;
                sj[q] = jj;
                v = tv;
                jj = tj;
            }
        }
    }
    for (int r = 0; r < kSeed; ++r) {
        double mv = sv[0];
        int mj = sj[0];
        wave_min_vj(mv, mj);
        if (mj == kNone) break;  // uniform: no lane has a finite bound left
        if (lane == 0) push_pair(pa, pb, cnt, i, mj);
        if (sj[0] == mj) {  // the lane holding it pops its head
#pragma unroll
            for (int q = 0; q + 1 < kSeed; ++q) {
                sv[q] = sv[q + 1];
                sj[q] = sj[q + 1];
            }
            sv[kSeed - 1] = WBX_INF_D;
            sj[kSeed - 1] = kNone;
        }
    }
}

// Row i's scan from column j0 to j1 with the running state (top, U): every
// pair with lb[j] < U(j) whose distance is not known is listed; a known value
// below U enters the k smallest.
__device__ __forceinline__ void scan_span(const double *__restrict__ drow,
                                          const double *__restrict__ lrow, int64_t i, int64_t j0,
                                          int64_t j1, int k, bool holds, double (&top)[kKMax],
                                          double &U, int32_t *__restrict__ pa,
                                          int32_t *__restrict__ pb,
                                          unsigned long long *__restrict__ cnt) {
    for (int64_t j = j0; j < j1; ++j) {
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

constexpr int kKnownCap = 512;  // known distances of a row gathered in LDS

// One wave per row: with D holding the known distances (NaN elsewhere),
// every pair with lb[j] < U(j) whose distance is not known is listed, U(j)
// the k-th smallest known value before j.  The bound must hold on the known
// pairs, else every pair with a finite bound is listed (the replay is then
// exact for any bound).  Lane l scans the columns of its chunk in order,
// starting from the state after the known values before the chunk (gathered
// in index order in LDS); a row with more than kKnownCap known values is
// scanned by one lane.
__global__ __launch_bounds__(64) void k_prune_scan(const double *__restrict__ D,
                                                   const double *__restrict__ lb, int64_t nx,
                                                   int64_t ny, int k, int32_t *__restrict__ pa,
                                                   int32_t *__restrict__ pb,
                                                   unsigned long long *__restrict__ cnt) {
    __shared__ double kv[kKnownCap];
    const int64_t i = blockIdx.x;
    if (i >= nx) return;
    const int lane = threadIdx.x;
    const double *drow = D + i * ny;
    const double *lrow = lb + i * ny;
    bool bad = false;
    for (int64_t j = lane; j < ny; j += 64) {
        const double v = drow[j];
        if (v == v && lrow[j] > v) bad = true;
    }
    const bool holds = __ballot(bad) == 0ull;
    const int64_t C = (ny + 63) / 64;
    const int64_t j0 = lane * C < ny ? lane * C : ny;
    const int64_t j1 = j0 + C < ny ? j0 + C : ny;
    int nk = 0;
    for (int64_t j = j0; j < j1; ++j) {
        const double v = drow[j];
        nk += v == v ? 1 : 0;
    }
    int inc = nk;  // inclusive prefix over the lanes
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(inc, d);
        if (lane >= d) inc += t;
    }
    const int total = __shfl(inc, 63);
    double top[kKMax];
#pragma unroll
    for (int q = 0; q < kKMax; ++q) top[q] = WBX_INF_D;
    double U = WBX_INF_D;
    if (total > kKnownCap) {  // uniform
        if (lane == 0) scan_span(drow, lrow, i, 0, ny, k, holds, top, U, pa, pb, cnt);
        return;
    }
    int o = inc - nk;
    for (int64_t j = j0; j < j1; ++j) {
        const double v = drow[j];
        if (v == v) kv[o++] = v;
    }
    __syncthreads();
    const int before = inc - nk;  // known values before this lane's chunk
    for (int e = 0; e < before; ++e) {
        const double v = kv[e];
        if (v < U) {
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
    scan_span(drow, lrow, i, j0, j1, k, holds, top, U, pa, pb, cnt);
}

}  // namespace

int prune_k_max() { return kKMax; }

int launch_prune_seed(const double *lb, int64_t nx, int64_t ny, int64_t k, int32_t *pa,
                      int32_t *pb, unsigned long long *cnt, hipStream_t st) {
    if (k > kKMax) return fail(WBX_EINVAL, "bound pruning serves k <= %d", kKMax);
    if (nx > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "bound pruning: too many rows");
    hipLaunchKernelGGL(k_prune_seed, dim3((unsigned)nx), dim3(64), 0, st, lb, nx, ny, k, pa, pb,
                       cnt);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

int launch_prune_scan(const double *D, const double *lb, int64_t nx, int64_t ny, int64_t k,
                      int32_t *pa, int32_t *pb, unsigned long long *cnt, hipStream_t st) {
    if (k > kKMax) return fail(WBX_EINVAL, "bound pruning serves k <= %d", kKMax);
    if (nx > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "bound pruning: too many rows");
    hipLaunchKernelGGL(k_prune_scan, dim3((unsigned)nx), dim3(64), 0, st, D, lb, nx, ny, (int)k,
                       pa, pb, cnt);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

}  // namespace wbx
